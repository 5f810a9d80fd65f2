"""Copy of a known size, for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes.

``rocprofv3 --pmc FETCH_SIZE -- python3 -m hnsw_amd.copy_calibration`` (and the same with
WRITE_SIZE): the device copy kernel moves exactly 1 GiB in and 1 GiB out per call, so
1,048,576 KiB divided by the counter's KiB per dispatch is the factor applied to other kernels.
"""

import torch


def main() -> None:
    n = 1 << 28  # float32 elements = 1 GiB
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied 3 x {n * 4} bytes")


if __name__ == "__main__":
    main()
