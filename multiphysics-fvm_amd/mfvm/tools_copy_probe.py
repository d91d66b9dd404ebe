"""Plain device copy of a known size (1 GiB, 3 times): the calibration kernel
for the FETCH_SIZE / WRITE_SIZE counters used in the traffic figures."""

import torch

if __name__ == "__main__":
    n = 1 << 27  # 128 Mi float64 = 1 GiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)       # __amd_rocclr_copyBuffer: reads 1 GiB, writes 1 GiB
    for _ in range(3):
        torch.add(a, a, out=b)  # elementwise kernel: reads 1 GiB (a twice, cached), writes 1 GiB
    torch.cuda.synchronize()
    print("copied", a.numel() * 8, "bytes x3", flush=True)
