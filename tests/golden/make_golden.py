"""Writes the committed fixtures under tests/golden/ (run in the build container,
where the reference tree and the reference-built checker libraries exist).

* sherman2.npz -- HB/sherman2 (the reference's only fixture,
  {AFIIM,PDFIIM,ACIFIM}/example/sherman2.mtx and sherman2_rhs1.mtx) as read by
  the reference's own reader (AFIIM/src/rw_matrix.cpp:62-309): CSR + rhs.
* afiim_sherman2.npz / acifim_sherman2.npz -- full factors produced by the
  reference's own afiim() / acifim() (compiled unmodified into oracle/_ref)
  with the README parameters (AFIIM/README:29 eps=9e-6; ACIFIM/README:31
  eps1=eps2=0.01).
* digests.json -- SHA-256 digests of the factor arrays the reference builds
  produce for more inputs (sherman2 at several tolerances, generated
  convection-diffusion and 27-point systems), plus digests of the generated
  matrices themselves.

Usage: python tests/golden/make_golden.py   (from the repository root)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "computational-mathematics_amd"), ROOT):
    sys.path.insert(0, p)

import cmamd  # noqa: E402
import oracle  # noqa: E402

REF = "/root/reference"


def digest_csr(m):
    h = hashlib.sha256()
    for a in (m.ptr, m.idx, m.val):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def digest_arr(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def digest_factors(F):
    d = {"G": digest_csr(F.G), "H": digest_csr(F.H), "idv": digest_arr(F.idv),
         "nnz": {"G": F.G.nnz, "H": F.H.nnz}}
    if F.L is not None:
        d["L"] = digest_csr(F.L)
        d["U"] = digest_csr(F.U)
        d["nnz"].update({"L": F.L.nnz, "U": F.U.nnz})
    return d


# Inputs of the digest table: name -> generator call
def inputs():
    sh = oracle.ref_read_coeff_matrix(os.path.join(REF, "AFIIM/example/sherman2.mtx"))
    return {
        "sherman2": sh,
        "convdiff2d_100_pe0.5_s1": cmamd.gen_convdiff(2, 100, 0.5, seed=1),
        "convdiff3d_16_pe5_s8": cmamd.gen_convdiff(3, 16, 5.0, seed=8),
        "stencil27_10_s9": cmamd.gen_stencil27(10, 0.05, 200, 600, seed=9),
    }


AFIIM_CASES = [("sherman2", 9e-6), ("sherman2", 0.01), ("sherman2", 0.1), ("sherman2", 0.0),
               ("convdiff2d_100_pe0.5_s1", 0.1), ("convdiff3d_16_pe5_s8", 0.05),
               ("stencil27_10_s9", 0.05)]
ACIFIM_CASES = [("sherman2", 0.01, 0.01, 200), ("sherman2", 0.1, 0.1, 200), ("sherman2", 0.001, 0.001, 5),
                ("convdiff3d_16_pe5_s8", 0.01, 0.1, 200), ("convdiff2d_100_pe0.5_s1", 0.01, 0.1, 200)]


def main():
    ins = inputs()
    sh = ins["sherman2"]
    rhs = oracle.ref_read_rhs_vector(os.path.join(REF, "AFIIM/example/sherman2_rhs1.mtx"))
    np.savez_compressed(os.path.join(HERE, "sherman2.npz"), ptr=sh.ptr, idx=sh.idx, val=sh.val, rhs=rhs)
    F = oracle.ref_afiim(sh, 9e-6)
    np.savez_compressed(os.path.join(HERE, "afiim_sherman2.npz"), gp=F.G.ptr, gi=F.G.idx, gv=F.G.val,
                        hp=F.H.ptr, hi=F.H.idx, hv=F.H.val, idv=F.idv, dtol=9e-6)
    rc, F = oracle.ref_acifim(sh, 0.01, 0.01)
    assert rc == 0
    np.savez_compressed(os.path.join(HERE, "acifim_sherman2.npz"), gp=F.G.ptr, gi=F.G.idx, gv=F.G.val,
                        hp=F.H.ptr, hi=F.H.idx, hv=F.H.val, lp=F.L.ptr, li=F.L.idx, lv=F.L.val,
                        up=F.U.ptr, ui=F.U.idx, uv=F.U.val, idv=F.idv, dtol1=0.01, dtol2=0.01)
    out = {"source": "reference afiim()/acifim() compiled unmodified into oracle/_ref (g++ -O3)",
           "inputs": {k: digest_csr(v) for k, v in ins.items()},
           "afiim": {}, "acifim": {}}
    for name, dtol in AFIIM_CASES:
        out["afiim"][f"{name}|{dtol!r}"] = digest_factors(oracle.ref_afiim(ins[name], dtol))
    for name, d1, d2, el in ACIFIM_CASES:
        rc, F = oracle.ref_acifim(ins[name], d1, d2, element_limit=el)
        assert rc == 0
        out["acifim"][f"{name}|{d1!r}|{d2!r}|{el}"] = digest_factors(F)
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
