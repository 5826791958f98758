"""Seeded scenes shared by the tests (same generators as the fixtures)."""
import sequoia_amd as sq

GOLDEN_SCENES = {
    "ico1280": lambda: sq.gen_icosphere(1280, 1, 0.1),
    "soup2000": lambda: sq.gen_soup(2000, 7, 0.01),
    "clus3000": lambda: sq.gen_clusters(3000, 16, 5),
    "ico69632": lambda: sq.gen_icosphere(69632, 1, 0.1),
}
