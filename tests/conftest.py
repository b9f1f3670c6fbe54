import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "cl-util_amd"), os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def score_crop():
    d = np.load(os.path.join(GOLDEN, "score_crop.npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


def crop_case(score_crop, name):
    s = score_crop
    H, W = s["source"].shape
    mask = np.unpackbits(s[name + "_mask"])[: H * W].reshape(H, W)
    return ((s["source"], s["sink"], s[name + "_right"], s[name + "_down"]),
            int(s[name + "_flow"][0]), mask)
