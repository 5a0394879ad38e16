"""Launch the H-matrix product alone, for kernel traces and counter collection.

    python -m qdfn.profile_hmat [--reps 50] [--crit 1.5] [--aca-tol 1e-6] [--no-calibration] [--sampled]
                                [--prec fp32|mixed|mixed8|bf16|fp64]

Builds the 20,000-triangle rough-fault H operator (storage --prec: FP32; FP32
with FP16 trailing low-rank components; BF16 low-rank factors; FP64) on the device
and runs `reps` products (phase 1: t = V x, phase 2: y = panels [x; t]).
Unless --no-calibration is given it first builds the 8,192-element
fault-network dense FP32 operator, whose build launches k_build, a plain
streaming kernel of known traffic (reads K_S and K_N in FP64: 1 GiB, writes E
and K_N in FP32: 512 MiB), against which the HBM counters (FETCH_SIZE /
WRITE_SIZE) are calibrated.  --sampled builds the operator from sampled rows
and columns (qdfn_hmat_build_tri_sampled) instead of the full assembly;
the storage is chosen by --prec.
"""
import argparse
import sys

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--crit", type=float, default=1.5)
    ap.add_argument("--aca-tol", type=float, default=1e-6)
    ap.add_argument("--no-calibration", action="store_true")
    ap.add_argument("--sampled", action="store_true")
    ap.add_argument("--prec", choices=("fp32", "mixed", "mixed8", "bf16", "fp64"), default="mixed")
    a = ap.parse_args()
    import qdfn
    from qdfn import configs
    if not a.no_calibration:
        cfg = configs.fault_network()
        op = qdfn.DenseOperator.from_rect_segments(cfg["seg"], cfg["mu"], cfg["nu"], qdfn.PREC_FP32)
        op.close()
        print("calibration operator built (k_build<float>: reads 1048576 KiB, writes 524288 KiB)", flush=True)
    m = configs.rough_fault()
    n = m["P1"].shape[0]
    P1, P2, P3, nrm, _, _, slip = qdfn.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]), n, m["rake"])
    ctr = (P1 + P2 + P3) / 3
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    st = qdfn.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=40, dist_diam_ratio=a.crit,
                             arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    geo = [x[p] for x in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    h = qdfn.HMatrix.build_tri(*geo, m["mu"], lam, st["range_SR"], st["admissible"], aca_rtol=a.aca_tol,
                               prec={"fp32": qdfn.PREC_FP32, "mixed": qdfn.PREC_MIXED, "mixed8": qdfn.PREC_MIXED8,
                                     "bf16": qdfn.PREC_BF16_FAR, "fp64": qdfn.PREC_FP64}[a.prec], sampled=a.sampled)
    print(f"build: {h.build_info()}", flush=True)
    info = h.info()
    t1, t2, tot = h.time_phases(a.reps)
    print(f"bytes by part: {h.bytes_by_part()}", flush=True)
    print(f"H product ({a.prec}, crit {a.crit}) n={n}: phase 1 {t1 * 1e3:.1f} us ({info['phase1_bytes'] / 1e6:.1f} MB), "
          f"phase 2 kernel {t2 * 1e3:.1f} us ({info['phase2_bytes'] / 1e6:.1f} MB, "
          f"{info['phase2_bytes'] / (t2 * 1e-3) / 1e9:.0f} GB/s), total {tot * 1e3:.1f} us "
          f"({info['bytes_per_product'] / 1e6:.1f} MB)", flush=True)
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
