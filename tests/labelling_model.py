"""A plain-Python model of the device's per-frame labelling (test
infrastructure): four-connected components of an ordered signal list,
numbered by first pixel, their exact integer moments, and the links to the
previous frame -- what label.hip hands the host.  Used on the CPU to test the
host-side grouping (sf_labeller_add_labelled) without a device."""
import numpy as np

import spotfinder_hip as sf


def label_frame(signal, width):
    idx = [int(i) for i in signal["index"]]
    val = [int(v) for v in signal["value"]]
    n = len(idx)
    rank = {i: r for r, i in enumerate(idx)}
    parent = list(range(n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    def unite(a, b):
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)

    for r, i in enumerate(idx):
        if i % width != 0 and (i - 1) in rank:
            unite(r, rank[i - 1])
        if i >= width and (i - width) in rank:
            unite(r, rank[i - width])
    roots = sorted({find(r) for r in range(n)})
    cid = {root: c for c, root in enumerate(roots)}
    comp_of = np.array([cid[find(r)] for r in range(n)], np.uint32)
    comps = np.zeros(len(roots), sf.COMPONENT)
    acc = [dict(n=0, x0=1 << 32, x1=0, y0=1 << 32, y1=0, peak=(0, 0), s=0,
                ss=0, sx=0, sy=0, sxx=0, syy=0) for _ in roots]
    for r, (i, v) in enumerate(zip(idx, val)):
        a = acc[comp_of[r]]
        x, y = i % width, i // width
        ax, ay = 2 * x + 1, 2 * y + 1
        a["n"] += 1
        a["x0"], a["x1"] = min(a["x0"], x), max(a["x1"], x + 1)
        a["y0"], a["y1"] = min(a["y0"], y), max(a["y1"], y + 1)
        a["peak"] = max(a["peak"], (v, 0xFFFFFFFF - i))
        a["s"] += v
        a["ss"] += v * v
        a["sx"] += v * ax
        a["sy"] += v * ay
        a["sxx"] += v * ax * ax
        a["syy"] += v * ay * ay
    mask = (1 << 64) - 1
    for c, root in enumerate(roots):
        a = acc[c]
        comps[c]["first"] = idx[root]
        for k in ("n", "x0", "x1", "y0", "y1"):
            comps[c][k] = a[k]
        comps[c]["peak"] = (a["peak"][0] << 32) | a["peak"][1]
        comps[c]["sum"] = a["s"]
        for k, key in (("ss", "sum_sq"), ("sx", "sum_x"), ("sy", "sum_y"),
                       ("sxx", "sum_xx"), ("syy", "sum_yy")):
            comps[c][key] = (a[k] & mask, a[k] >> 64)
    return comp_of, comps, rank


def links_between(signal, comp_of, prev_rank, prev_comp_of):
    pairs = set()
    for r, i in enumerate(int(x) for x in signal["index"]):
        if i in prev_rank:
            pairs.add((int(comp_of[r]), int(prev_comp_of[prev_rank[i]])))
    return np.array(sorted(pairs), np.uint32).reshape(-1, 2)


class ModelFeeder:
    """Feeds frames to an sf.Labeller through the device's contract,
    computing the device's outputs with the model above."""

    def __init__(self, labeller, width):
        self.lab = labeller
        self.width = width
        self.prev = None  # (frame, rank, comp_of)

    def add(self, frame, signal):
        signal = np.ascontiguousarray(signal, sf.SIGNAL_PIXEL)
        comp_of, comps, rank = label_frame(signal, self.width)
        links = np.zeros((0, 2), np.uint32)
        linked = self.prev is not None and self.prev[0] + 1 == frame
        if linked:
            links = links_between(signal, comp_of, self.prev[1], self.prev[2])
        self.lab.add_labelled(frame, signal, comp_of, comps, links, linked)
        self.prev = (frame, rank, comp_of)
