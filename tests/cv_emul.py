"""Reference evaluations of packed control-variate networks (test support).

`lane_forward` replays the fused kernel's data flow lane by lane: the MFMA
operand and accumulator layouts of v_mfma_f32_32x32x16_bf16 (lane l holds
row / column l % 32 with the 8 K positions of half l / 32; accumulator r of
lane l is row 8 (r / 4) + 4 (l / 32) + r % 4), the accumulator-to-operand
reuse between layers and the bf16 high/low split of the inputs.
`matrix_forward` computes the same bf16 network as plain matrix products on
the padded matrices.  They agree only if the packed word order, the unit
order u(c, h, q) and the kernel's data flow fit together.
"""
import numpy as np
import torch


def bf16(a):
    """Round to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def words_to_float(words):
    return (words.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _mfma(A, B, C):
    """32x32 tile product from lane fragments A[64, 8], B[64, 8]."""
    Am = np.concatenate([A[:32], A[32:]], axis=1)          # [32 rows, (h, q)]
    Bm = np.concatenate([B[:32], B[32:]], axis=1)          # [32 cols, (h, q)]
    return C + Am @ Bm.T


def _regs(Cm):
    """Accumulator registers [64 lanes, 16] of a 32x32 tile."""
    lane = np.arange(64)[:, None]
    r = np.arange(16)[None, :]
    return Cm[8 * (r // 4) + 4 * (lane // 32) + r % 4, lane % 32]


def lane_forward(words, hidden, F):
    """Outputs [64, 8] of the packed network for 64 feature rows F[64, 8]."""
    blk = words_to_float(words).reshape(-1, 64, 8)
    hi = bf16(F)
    lo = bf16(F - hi)
    lane = np.arange(64)
    ins = []
    for T in range(2):
        path = 32 * T + lane % 32
        ins.append(np.where((lane // 32 == 0)[:, None], hi[path], lo[path]))
    acts = []
    for T in range(2):
        chunks = []
        for m in range(2):
            R = _regs(_mfma(blk[m], ins[T], np.zeros((32, 32))))
            chunks += [bf16(np.maximum(R[:, 8 * s:8 * s + 8], 0.0)) for s in range(2)]
        acts.append(chunks)
    b = 2
    for _ in range(1, hidden):
        for T in range(2):
            chunks = []
            for m in range(2):
                Cm = np.zeros((32, 32))
                for c in range(4):
                    Cm = _mfma(blk[b + 4 * m + c], acts[T][c], Cm)
                R = _regs(Cm)
                chunks += [bf16(np.maximum(R[:, 8 * s:8 * s + 8], 0.0)) for s in range(2)]
            acts[T] = chunks
        b += 8
    out = np.zeros((64, 8))
    for T in range(2):
        Cm = np.zeros((32, 32))
        for c in range(4):
            Cm = _mfma(blk[b + c], acts[T][c], Cm)
        out[32 * T:32 * T + 32] = Cm[:8].T
    return out


def matrix_forward(mats, F):
    """The bf16 network as matrix products: bf16 weights, inputs split into
    bf16 high and low parts, activations rounded to bf16 after the ReLU."""
    Wb = [bf16(M) for M in mats]
    hi = bf16(F)
    lo = bf16(F - hi)
    a = bf16(np.maximum((hi + lo) @ Wb[0].T, 0.0))
    for W in Wb[1:-1]:
        a = bf16(np.maximum(a @ W.T, 0.0))
    return (a @ Wb[-1].T)[:, :8]


def torch_forward(mats, F):
    """matrix_forward in torch (float32 products) for device-side checks;
    F is a float32 tensor [n, 8]."""
    Wb = [torch.tensor(M, dtype=torch.float32, device=F.device).to(torch.bfloat16)
          .to(torch.float32) for M in mats]
    hi = F.to(torch.bfloat16).to(torch.float32)
    lo = (F - hi).to(torch.bfloat16).to(torch.float32)
    a = torch.relu((hi + lo) @ Wb[0].T).to(torch.bfloat16).to(torch.float32)
    for W in Wb[1:-1]:
        a = torch.relu(a @ W.T).to(torch.bfloat16).to(torch.float32)
    return (a @ Wb[-1].T)[:, :8]
