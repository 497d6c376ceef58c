"""Writes the golden fixtures of the reference's own test data.

Reads (in this container only) the MatrixMarket / text files that the
reference's tests hold and stores them as compact .npz data:
  set_C.npz      test/PETSc/data/set_C/{matrix,rhs,sol,x0}.txt
                 (65x65-node 2D Q1 Poisson system, 4225 rows, 37249 entries)
  equation1.npz  test/LinearAlgebra/data/Equation1.{A,rhs,x0,sol}.txt
  mm/            the small MatrixMarket files of the reference's tests, as they
                 are (test/LinearAlgebra/data/Equation1.A.txt,
                 test/AMG/data/A3.mtx, test/AMG/data/A4.mtx)
Run:  python tests/golden/make_fixtures.py /root/reference
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_mm(path):
    rows, cols, vals = [], [], []
    with open(path) as f:
        header = f.readline()
        assert header.startswith("%%MatrixMarket matrix coordinate real general"), header
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        n, m, nnz = map(int, line.split())
        for line in f:
            if not line.strip():
                continue
            i, j, v = line.split()
            rows.append(int(i) - 1)
            cols.append(int(j) - 1)
            vals.append(float(v))
    assert len(vals) == nnz
    return n, m, np.array(rows, np.int32), np.array(cols, np.int32), np.array(vals)


def read_vec(path):
    return np.array([float(t) for t in open(path).read().split()])


def main(ref):
    d = os.path.join(ref, "test/PETSc/data/set_C")
    n, m, r, c, v = read_mm(os.path.join(d, "matrix.txt"))
    np.savez_compressed(os.path.join(HERE, "set_C.npz"), n=n, rows=r, cols=c, vals=v,
                        rhs=read_vec(os.path.join(d, "rhs.txt")), sol=read_vec(os.path.join(d, "sol.txt")),
                        x0=read_vec(os.path.join(d, "x0.txt")))
    d = os.path.join(ref, "test/LinearAlgebra/data")
    n, m, r, c, v = read_mm(os.path.join(d, "Equation1.A.txt"))
    np.savez_compressed(os.path.join(HERE, "equation1.npz"), n=n, rows=r, cols=c, vals=v,
                        rhs=read_vec(os.path.join(d, "Equation1.rhs.txt")),
                        sol=read_vec(os.path.join(d, "Equation1.sol.txt")),
                        x0=read_vec(os.path.join(d, "Equation1.x0.txt")))


def copy_mm(ref):
    import shutil
    os.makedirs(os.path.join(HERE, "mm"), exist_ok=True)
    for src in ("test/LinearAlgebra/data/Equation1.A.txt", "test/LinearAlgebra/data/Equation1.rhs.txt",
                "test/LinearAlgebra/data/Equation1.sol.txt", "test/LinearAlgebra/data/Equation1.x0.txt",
                "test/AMG/data/A3.mtx", "test/AMG/data/A4.mtx"):
        shutil.copyfile(os.path.join(ref, src), os.path.join(HERE, "mm", os.path.basename(src)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    copy_mm(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
