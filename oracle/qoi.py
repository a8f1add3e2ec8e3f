"""QoI restatement (numpy) -- TEST INFRASTRUCTURE ONLY.

Follows 3DgrainMPI/QOI.cpp: line QoIs (tip scan, per-grain areas, :38-95);
per-grain volume over interior points with
alpha > 0 in first-seen (k, j, i) scan order (:198-231); 3-D junctions: solid
interior points whose 19-point neighbourhood holds no liquid and exactly 4
distinct ids with the largest count <= 10 (:233-311); ids compared as sorted
sets (the reference's order comes from unordered_map iteration).
2-D junctions of the line configuration on the tip planes of the archived
alpha (QOILine::searchJunctionsOnImage, :103-160): interior points of each
frame's tip plane whose 3x3 window holds at least 3 distinct values (liquid
0 included) with the largest count <= 5; the values are written ascending
(the reference's order comes from unordered_map iteration), at most 5 of
them (the reference's row has room for 5 and writes a 6th into the next
row), and rows past a frame's share of node_region are skipped (the
reference writes into the next frame's share).
"""
import numpy as np


def volume(alpha, fnx, fny, fnz):
    a = np.asarray(alpha).reshape(fnz, fny, fnx)[1:-1, 1:-1, 1:-1].ravel()
    a = a[a > 0]
    ids, first = np.unique(a, return_index=True)
    order = np.argsort(first)
    counts = np.bincount(np.searchsorted(ids, a), minlength=ids.size)
    return ids[order].tolist(), counts[order].tolist()


def line_qois(alpha, fnx, fny, fnz, cur_tip, num_grains):
    """QOI::calculateLineQoIs (3DgrainMPI/QOI.cpp:38-95) without the loss-area
    term: step the tip up to the first plane whose interior holds liquid
    (alpha 0), minus one (:47-67; bounded at fnz - 2), then per grain id
    1..num_grains the interior points in all planes (total) and above the tip
    (extra) (:71-90)."""
    A = np.asarray(alpha).reshape(fnz, fny, fnx)
    core = A[:, 1:-1, 1:-1]
    tip = cur_tip
    while True:
        tip += 1
        if tip >= fnz - 1 or (core[tip] == 0).any():
            break
    tip -= 1
    inner = A[1:-1, 1:-1, 1:-1]
    ks = np.arange(1, fnz - 1)[:, None, None] * np.ones_like(inner)
    ok = (inner > 0) & (inner <= num_grains)
    total = np.bincount(inner[ok] - 1, minlength=num_grains)[:num_grains]
    up = ok & (ks > tip)
    extra = np.bincount(inner[up] - 1, minlength=num_grains)[:num_grains]
    return tip, total, extra


def junctions3d(alpha, fnx, fny, fnz):
    A = np.asarray(alpha).reshape(fnz, fny, fnx)
    offs = [(dk, dj, di) for dk in (-1, 0, 1) for dj in (-1, 0, 1) for di in (-1, 0, 1)
            if di * di + dj * dj + dk * dk <= 2]
    core = A[1:-1, 1:-1, 1:-1]
    nb = np.stack([A[1 + dk:fnz - 1 + dk, 1 + dj:fny - 1 + dj, 1 + di:fnx - 1 + di] for dk, dj, di in offs])
    out = []
    cand = np.argwhere((core != 0) & ~(nb == 0).any(0))
    for k, j, i in cand:
        v = nb[:, k, j, i]
        ids, cnt = np.unique(v, return_counts=True)
        if ids.size == 4 and cnt.max() <= 10:
            out.append((int(i + 1), int(j + 1), int(k + 1), int(cnt.max()), tuple(int(x) for x in ids)))
    return out


def line_junctions(alpha_full, fnx, fny, tip_y, num_nodes, nfeat=8, repeated=5):
    """QOILine::searchJunctionsOnImage (3DgrainMPI/QOI.cpp:103-160) on the
    full-height alpha (fnz_full planes of fnx * fny, halo included) at the
    frames' tip planes tip_y.  Returns (cross_sec, node_region): the tip
    planes one after another, and rows of nfeat ints (i, j, largest count,
    then the ids) in each frame's share of repeated * num_nodes rows, -1
    elsewhere."""
    A = np.asarray(alpha_full).reshape(-1, fny, fnx)
    per_frame = repeated * num_nodes * nfeat
    nodes = np.full(len(tip_y) * per_frame, -1, np.int32)
    cross = []
    for kt, tip in enumerate(tip_y):
        P = A[int(tip)]
        cross.append(P.ravel().astype(np.int32))
        cnt = 0
        for j in range(1, fny - 1):
            for i in range(1, fnx - 1):
                ids, c = np.unique(P[j - 1:j + 2, i - 1:i + 2], return_counts=True)
                if ids.size >= 3 and c.max() <= 5:
                    base = per_frame * kt + cnt * nfeat
                    if base + nfeat > per_frame * (kt + 1):
                        continue
                    nodes[base:base + 3] = (i, j, int(c.max()))
                    for q, v in enumerate(ids[:nfeat - 3]):
                        nodes[base + 3 + q] = int(v)
                    cnt += 1
    return np.concatenate(cross) if cross else np.zeros(0, np.int32), nodes
