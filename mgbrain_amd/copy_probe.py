"""Device-to-device copy of a known size, the reference point for converting
rocprofv3 FETCH_SIZE / WRITE_SIZE counters into HBM bytes:
    rocprofv3 --pmc FETCH_SIZE -- python3 -m mgbrain_amd.copy_probe
The copy moves exactly NBYTES read + NBYTES written per launch."""
import sys

NBYTES = 1 << 30


def main():
    import torch  # plumbing only: device buffers and a copy kernel

    a = torch.empty(NBYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {NBYTES} bytes x 3", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
