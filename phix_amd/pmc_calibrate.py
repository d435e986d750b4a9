"""Known-size kernels for calibrating rocprofv3's FETCH_SIZE and WRITE_SIZE
counters: 3 launches each of a read-only, a write-only and a copy kernel
(phx_hbm_probe kinds 0, 1, 2) over 1 GiB FP32 buffers. Run it under
`rocprofv3 --pmc FETCH_SIZE` (and again under WRITE_SIZE); each kernel moves
exactly 1 GiB per direction it touches.

    python -m phix_amd.pmc_calibrate
"""
import sys


def main() -> int:
    import torch
    from phix_amd import capi
    n = (1 << 30) // 4
    a = torch.zeros(n, dtype=torch.float32, device="cuda:0")
    b = torch.zeros(n, dtype=torch.float32, device="cuda:0")
    sh = capi.stream_handle()
    P = capi.ptr
    for _ in range(3):
        capi.call("phx_hbm_probe", 0, P(a), P(b), None, None, None, n, 0, 0, 8192, sh)
        capi.call("phx_hbm_probe", 1, P(a), None, None, None, None, n, 0, 0, 8192, sh)
        capi.call("phx_hbm_probe", 2, P(a), P(b), None, None, None, n, 0, 0, 0, sh)
    torch.cuda.synchronize()
    print("calibration kernels done: read, write, copy of 1 GiB, 3 launches each", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
