"""The GrowCut oracle (oracle/growcut_oracle.py) pinned on hand-worked
cases: a label travels one pixel per step through a flat image, the colour
weight g(X) = 1 - X / sqrt(3) at a known distance, the neighbour order on
ties, and a planted mistake (the order reversed) caught by the tie case."""
import numpy as np

import growcut_oracle as GC


def test_flat_image_propagates_one_pixel_per_step():
    img = np.full((1, 6, 3), 77, np.uint8)
    for k in range(1, 6):
        lab, st, _ = GC.run(img, [(np.array([0]), 3)], steps=k)
        assert list(lab[0]) == [3] * (k + 1) + [0] * (5 - k)
        assert (st[0, : k + 1] == 1.0).all()
    lab, st, changed = GC.run(img, [(np.array([0]), 3)])
    assert (lab == 3).all() and changed == 5


def test_weight_at_a_colour_edge():
    # black | red: X = |(1, 0, 0, 0)| = 1, g = 1 - 1 / sqrt(3)
    img = np.zeros((1, 2, 3), np.uint8)
    img[0, 1, 0] = 255
    lab, st, _ = GC.run(img, [(np.array([0]), 1)], steps=1)
    g = np.float32(1.0) - np.float32(1.0) / np.float32(1.7320508075688772)
    assert lab[0, 1] == 1 and st[0, 1] == g
    assert abs(float(g) - (1 - 1 / 3 ** 0.5)) < 1e-7
    # the other weights at x = 1 / sqrt(3)
    x = np.float32(1.0) / np.float32(1.7320508075688772)
    for kind, want in ((1, 1 - x * x), (2, 1 - x ** 3), (4, 1 - np.sqrt(x))):
        assert abs(float(GC.weight(kind, np.float32(1.0))) - float(want)) < 1e-6


def test_neighbour_order_on_ties_and_planted_mistake():
    # a pixel between an up and a down seed of equal strength and colour:
    # up attacks first and a tie does not take over, so up's label wins
    img = np.full((3, 1, 3), 10, np.uint8)
    lab, _, _ = GC.run(img, [(np.array([0]), 1), (np.array([2]), 2)], steps=1)
    assert lab[1, 0] == 1
    # the same automaton with the order reversed picks the other label
    c = GC.colours(img)
    l0 = np.array([[1], [0], [2]], np.uint8)
    s0 = np.array([[1.0], [0.0], [1.0]], np.float32)
    defence, label = s0.copy(), l0.copy()
    for dy in (1, -1):  # down first: a planted mistake
        a = GC.weight(0, np.float32(0.0)) * s0[1 + dy, 0]
        if a > defence[1, 0]:
            defence[1, 0], label[1, 0] = a, l0[1 + dy, 0]
    assert label[1, 0] == 2 != GC.step(c, l0, s0)[0][1, 0]
