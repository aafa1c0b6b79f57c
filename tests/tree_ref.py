"""Independent float64 model of crnlib's tree clusterizer, written from the
algorithm's description (crn_tree_clusterizer.h:89-171, 265-570) with numpy
array operations in plain position order.  Used only to pin the oracle's
structure (seeds, principal axis, Lloyd passes, queue order, leaf order) on
continuous data, where float32 and float64 rounding do not change any
decision."""
import heapq

import numpy as np


def tree_codebook_f64(vectors, weights, max_splits):
    v = np.asarray(vectors, np.float64)
    w = np.asarray(weights, np.float64)
    wi = np.asarray(weights, np.int64)
    wv = v * w[:, None]
    wd = (v * v).sum(1) * w

    def stats(idx):
        W = wi[idx].sum()
        s = wv[idx].sum(0)
        return s / W, W, wd[idx].sum() - s @ s / W

    nodes = []
    c, W, var = stats(np.arange(len(v)))
    nodes.append(dict(idx=np.arange(len(v)), c=c, w=W, var=var, left=-1))
    heap = [(-var, 0)]
    splits = 1
    while splits < max_splits and heap:
        negvar, ni = heap[0]
        nd = nodes[ni]
        if nd["var"] <= 0 or len(nd["idx"]) == 1:
            break
        heapq.heappop(heap)
        splits += 1
        idx = nd["idx"]
        X = v[idx]
        f = X[np.argmax(((X - nd["c"]) ** 2).sum(1))]
        o = X[np.argmax(((X - f) ** 2).sum(1))]
        lc, rc = (f + nd["c"]) / 2, (o + nd["c"]) / 2
        if len(idx) > 2:
            D = X - nd["c"]
            cov = (D * w[idx, None]).T @ D / nd["w"]
            axis = np.ones(v.shape[1])
            for _ in range(10):
                x = cov @ axis
                if x.max() != 0:
                    x = x / x.max()
                axis = x
            axis = axis / np.sqrt(axis @ axis)
            left = D @ axis < 0
            if w[idx][left].sum() > 0 and w[idx][~left].sum() > 0:
                lc = wv[idx][left].sum(0) / w[idx][left].sum()
                rc = wv[idx][~left].sum(0) / w[idx][~left].sum()
        prev = 1e10
        unsplittable = False
        for _ in range(1024):
            left = ((X - lc) ** 2).sum(1) < ((X - rc) ** 2).sum(1)
            if wi[idx][left].sum() == 0 or wi[idx][~left].sum() == 0:
                unsplittable = True
                break
            (lc, lw, lvar), (rc, rw, rvar) = stats(idx[left]), stats(idx[~left])
            total = lvar + rvar
            if total < 1e-5 or (prev - total) / total < 1e-5:
                break
            prev = total
        if unsplittable:
            continue
        li = len(nodes)
        nd["left"] = li
        nodes.append(dict(idx=idx[left], c=lc, w=lw, var=lvar, left=-1))
        nodes.append(dict(idx=idx[~left], c=rc, w=rw, var=rvar, left=-1))
        heapq.heappush(heap, (-lvar, li))
        heapq.heappush(heap, (-rvar, li + 1))
    leaf = np.zeros(len(v), np.int64)
    cb = []
    for nd in nodes:
        if nd["left"] != -1:
            continue
        leaf[nd["idx"]] = len(cb)
        cb.append(nd["c"])
    return np.array(cb), leaf
