"""MatrixMarket test inputs written on the fly (no downloads).

mycielskian(k) builds the Mycielski graphs of the SuiteSparse group
"Mycielski" (mycielskian2 ... mycielskianK in the reference's sweep list,
testbmv.sh:5-19): M2 = K2, and M(k+1) adds a twin u_i for every vertex v_i
(u_i adjacent to the neighbours of v_i) and one vertex w adjacent to every
twin. Vertices are numbered v_1..v_n, u_1..u_n, w.
"""
import numpy as np


def mycielskian(k):
    """(n, edges i < j as an int array [E, 2]) of the Mycielski graph M_k, k >= 2."""
    n, edges = 2, [(0, 1)]
    for _ in range(k - 2):
        new = list(edges)
        for i, j in edges:  # twin of i joins the neighbours of i, and vice versa
            new.append((j, n + i))
            new.append((i, n + j))
        w = 2 * n
        new.extend((n + i, w) for i in range(n))
        n, edges = 2 * n + 1, new
    e = np.array(sorted((min(a, b), max(a, b)) for a, b in edges), dtype=np.int64)
    return n, e


def write_mtx(path, M, N, entries, sym="general"):
    """Pattern coordinate file with 1-based entries (i, j)."""
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate pattern {sym}\n")
        f.write(f"{M} {N} {len(entries)}\n")
        for i, j in entries:
            f.write(f"{int(i) + 1} {int(j) + 1}\n")


def write_mycielskian(path, k):
    """Symmetric pattern file (lower triangle, as SuiteSparse stores it)."""
    n, e = mycielskian(k)
    write_mtx(path, n, n, [(j, i) for i, j in e], sym="symmetric")
    return n, len(e)
