"""Launch the fused coupling kernel alone, for counter collection.

    python -m qdfn.profile_coupling [--prec fp32|fp64] [--reps 50] [--copy-mib 512]

Builds the 8,192-element fault-network operator on the device and runs `reps`
coupling products.  The operator build itself launches k_build, a plain
streaming kernel of known traffic (reads K_S and K_N in FP64, writes E and
K_N at the stored precision), against which the HBM counters (FETCH_SIZE /
WRITE_SIZE) of the coupling kernel are calibrated.  --copy-mib adds a torch
elementwise copy of that size as a second calibration kernel (it does not run
under rocprofv3 counter collection on every image).
"""
import argparse
import sys

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--prec", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--copy-mib", type=int, default=0)
    a = ap.parse_args()
    import qdfn
    from qdfn import configs
    cfg = configs.fault_network()
    prec = qdfn.PREC_FP32 if a.prec == "fp32" else qdfn.PREC_FP64
    op = qdfn.DenseOperator.from_rect_segments(cfg["seg"], cfg["mu"], cfg["nu"], prec)
    op.matvec(np.ones(op.n))
    ms = op.time_coupling(a.reps)
    gbs = op.bytes_per_product() / (ms * 1e-3) / 1e9
    print(f"coupling {a.prec}: {ms * 1e3:.1f} us/launch, {op.bytes_per_product() / 1e6:.1f} MB algorithmic, "
          f"{gbs:.0f} GB/s", flush=True)
    if a.copy_mib > 0:
        import torch
        x = torch.empty(a.copy_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.mul(x, 1.0, out=y)  # elementwise kernel: reads and writes copy_mib MiB
        torch.cuda.synchronize()
        print(f"copy {a.copy_mib} MiB done", flush=True)
    op.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
