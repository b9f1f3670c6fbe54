"""CPU oracle of GrowCut (float32 numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.

Restates the reference's automaton (evolveVonNeumann growcut.cl:50-179,
GrowCut.evolve GrowCut.py:126-153) as a synchronous update of the whole
image: every pixel is attacked by its neighbours up, down, left, right, in
that order; attack = g(|c_p - c_q|) * strength_q with colours c / 255 per
channel (alpha included), |.| = sqrt(((dr^2 + dg^2) + db^2) + da^2), g of
x = |.| / sqrt(3); an attack above the current defence takes over the
neighbour's label and strength.  Pixels outside the image never attack.
float32 throughout, each operation rounded once in this order.
"""
import numpy as np

SQRT3 = np.float32(1.7320508075688772)


def colours(img):
    a = np.asarray(img, np.uint8)
    c = a.astype(np.float32) / np.float32(255.0)
    if a.shape[2] == 3:
        c = np.concatenate([c, np.zeros(a.shape[:2] + (1,), np.float32)], axis=2)
    return c


def weight(kind, X):
    x = X / SQRT3
    one = np.float32(1.0)
    if kind == 1:
        return one - x * x
    if kind == 2:
        return one - (x * x) * x
    if kind == 3:
        return one - x * np.sqrt(x)
    if kind == 4:
        return one - np.sqrt(x)
    return one - x


def step(c, lab, strength, kind=0):
    H, W = lab.shape
    defence = strength.copy()
    label = lab.copy()
    for dy, dx in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        ys = slice(max(0, -dy), H - max(0, dy))
        xs = slice(max(0, -dx), W - max(0, dx))
        qy = slice(max(0, dy), H - max(0, -dy))
        qx = slice(max(0, dx), W - max(0, -dx))
        d = c[ys, xs] - c[qy, qx]
        dist = np.sqrt(((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
                       + d[..., 3] * d[..., 3])
        a = weight(kind, dist) * strength[qy, qx]
        win = a > defence[ys, xs]
        defence[ys, xs] = np.where(win, a, defence[ys, xs])
        label[ys, xs] = np.where(win, lab[qy, qx], label[ys, xs])
    return label, defence


def run(img, seeds, steps=None, kind=0):
    """seeds: list of (flat indices, label).  Returns (labels, strengths,
    steps that changed something); steps=None runs to the fixed point."""
    c = colours(img)
    H, W = c.shape[:2]
    lab = np.zeros((H, W), np.uint8)
    strength = np.zeros((H, W), np.float32)
    for idx, l in seeds:
        lab.reshape(-1)[idx] = l
        strength.reshape(-1)[idx] = 1.0
    changed = 0
    k = 0
    while steps is None or k < steps:
        nl, ns = step(c, lab, strength, kind)
        k += 1
        if (nl == lab).all() and (ns == strength).all():
            if steps is None:
                break
        else:
            changed = k
        lab, strength = nl, ns
    return lab, strength, changed
