"""Shared helpers for the test suite."""
from __future__ import annotations

import numpy as np

TREE_KEYS = ("root", "child", "count", "start", "level", "sorted")


def ic(workload: int, n: int) -> dict:
    import stellarsolver_amd as S

    return S.generate_ic(S.ic_params(workload), n)


def tree_equal(a: dict, b: dict, keys=TREE_KEYS) -> bool:
    if a["ncells"] != b["ncells"]:
        return False
    for k in keys:
        if a.get(k) is None or b.get(k) is None:
            continue
        if not np.array_equal(np.asarray(a[k]), np.asarray(b[k])):
            return False
    return True


def first_tree_difference(a: dict, b: dict, keys=TREE_KEYS) -> str:
    if a["ncells"] != b["ncells"]:
        return f"ncells {a['ncells']} != {b['ncells']}"
    for k in keys:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if x.shape != y.shape:
            return f"{k}: shape {x.shape} != {y.shape}"
        d = np.flatnonzero(x != y)
        if d.size:
            i = int(d[0])
            return f"{k}[{i}]: {x[i]} != {y[i]} ({d.size} differ)"
    return "equal"


def rms_rel(ax, ay, rx, ry) -> float:
    num = (np.asarray(ax, np.float64) - rx) ** 2 + (np.asarray(ay, np.float64) - ry) ** 2
    den = np.asarray(rx, np.float64) ** 2 + np.asarray(ry, np.float64) ** 2
    return float(np.sqrt(np.mean(num / den)))
