"""Synthetic R1CS workloads (harness side, in the role of exp_cell's builders).

The reference benchmarks its own x*x=y circuit (snark-ring/examples/
exp_cell.rs:38-65).  The configurations measured here use random R1CS
instead, drawn from an instance seed with SplitMix64 (not `rand`):

* square n x n matrices, column 0 is the constant 1;
* columns [1, n/2) are uniform limbs of Z/2^k (the inputs);
* row i < n/2 gets A and B rows with 1..3 nonzeros (A: 1..4 for the MIMD
  serving mix) at uniform columns in [0, n/2) with uniform coefficients,
  and C[i, n/2+i] = 1 with the auxiliary z[n/2+i] = (Az)_i * (Bz)_i mod 2^k;
* rows >= n/2 are empty.

Matrices are COO triples (rows u32, cols u32, vals u64 reduced mod 2^k);
witnesses are u64 arrays of limbs.  Duplicated (row, col) entries are legal
and sum, as in the reference's COO matvecs (reduction/src/matrix.rs:11-18).
"""
import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)


def _mix(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def stream(seed, tag, count):
    """`count` SplitMix64 outputs for (seed, tag)."""
    with np.errstate(over="ignore"):
        base = _mix(np.uint64(seed & (2**64 - 1)) ^ (np.uint64(tag) * np.uint64(0xD1B54A32D192ED03)))
        idx = np.arange(1, count + 1, dtype=np.uint64)
        return _mix(base + idx * _G)


def mask(k):
    return np.uint64((1 << k) - 1) if k < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)


def random_matrices(n, k, seed, max_a=3, max_b=3):
    """A, B, C of the half-constrained random layout (n a power of two >= 4)."""
    half = n // 2
    msk = mask(k)
    out = []
    for tag, mx in ((1, max_a), (2, max_b)):
        cnt = (stream(seed, tag, half) % np.uint64(mx)).astype(np.int64) + 1
        rows = np.repeat(np.arange(half, dtype=np.uint32), cnt)
        nnz = len(rows)
        cols = (stream(seed, tag + 10, nnz) % np.uint64(half)).astype(np.uint32)
        vals = stream(seed, tag + 20, nnz) & msk
        out.append((rows, cols, vals))
    rows = np.arange(half, dtype=np.uint32)
    out.append((rows, (rows + half).astype(np.uint32), np.ones(half, np.uint64)))
    return tuple(out)


def _matvec(m, z, n, k):
    r, c, v = m
    msk = mask(k)
    with np.errstate(over="ignore"):
        prod = (v * z[c]) & msk
    out = np.zeros(n, np.uint64)
    # wrapping accumulation: reduce per row with np.add.at on uint64 (wraps mod 2^64)
    np.add.at(out, r, prod)
    return out & msk


def random_witness(n, k, A, B, seed, tag=0):
    """Satisfying witness for (A, B, C) from random_matrices."""
    half = n // 2
    msk = mask(k)
    z = np.zeros(n, np.uint64)
    z[0] = 1
    z[1:half] = stream(seed, 100 + tag, half - 1) & msk
    az = _matvec(A, z, half, k)
    bz = _matvec(B, z, half, k)
    with np.errstate(over="ignore"):
        z[half:] = (az * bz) & msk
    return z


def is_satisfied(A, B, Cm, z, n, k):
    msk = mask(k)
    az, bz, cz = (_matvec(m, z, n, k) for m in (A, B, Cm))
    with np.errstate(over="ignore"):
        return bool(np.all(((az * bz) & msk) == cz))


def general_instance(log_n, k, seed):
    """Config-1 style General instance: n = 2^log_n (log_blocks = log_n - 5)."""
    n = 1 << log_n
    A, B, Cm = random_matrices(n, k, seed)
    w = random_witness(n, k, A, B, seed)
    return A, B, Cm, w


def simd_instance(log, k, seed):
    """Config-0/2 style SIMD batch: shared n x n matrices (n = 2^log) and 32
    witnesses with independent inputs; returns (A, B, C, wit[32, n])."""
    n = 1 << log
    A, B, Cm = random_matrices(n, k, seed)
    wit = np.stack([random_witness(n, k, A, B, seed, tag=1 + s) for s in range(32)])
    return A, B, Cm, wit


def mimd_instance(log, k, seed, max_a=3):
    """MIMD batch: 32 independent instances, each with its own n x n matrices
    (n = 2^log) and satisfying witness.  Returns (As, Bs, Cs, wit[32, n])
    with As[s] = (rows, cols, vals) of instance s."""
    n = 1 << log
    seeds = stream(seed, 300, 32)
    As, Bs, Cs, ws = [], [], [], []
    for s in range(32):
        sd = int(seeds[s])
        A, B, Cm = random_matrices(n, k, sd, max_a=max_a)
        As.append(A)
        Bs.append(B)
        Cs.append(Cm)
        ws.append(random_witness(n, k, A, B, sd))
    return As, Bs, Cs, np.stack(ws)


def set_arrays(ms):
    """The 32 COO matrices of one MIMD set concatenated: (off[33], rows,
    cols, vals) with instance s at entries off[s] .. off[s+1]-1."""
    off = np.zeros(33, np.uint64)
    off[1:] = np.cumsum([len(m[0]) for m in ms])
    rows = np.ascontiguousarray(np.concatenate([m[0] for m in ms]), np.uint32)
    cols = np.ascontiguousarray(np.concatenate([m[1] for m in ms]), np.uint32)
    vals = np.ascontiguousarray(np.concatenate([m[2] for m in ms]), np.uint64)
    return off, rows, cols, vals


def png_instance(log_bytes, image="example"):
    """The PNG-filter batch of png_filter.rs (snark-ring/examples/png_filter.rs):
    32 grayscale images of b = 2^(log_bytes-5) bytes, W = min(256, b) columns,
    row r filtered with Sub (left neighbour) when r is even and Up (the row
    above) when odd (lines 84-103).  The ring-u8 circuit (build_ring, lines
    126-163): row i says Orig[i] * 1 = Filt[i] + Orig[pred(i)] over the
    witness w = (1, Orig[0], Filt[0], Orig[1], Filt[1], ...), n = b rows and
    m = 2^ceil_log2(1 + 2b) columns (lines 270-273).  Returns
    (log_n, log_m, A, B, C, wit[32, m]) with k = 8.

    image="example": raw bytes from the example's own SplitMix64 byte
    generator (lines 67-73, seed (slot << 40) ^ i * 0x10000001).
    image="gradient": each row a linear gradient (start byte and slope in
    [-4, 4] per row and image) plus uniform noise in [-16, 16], clamped to
    [0, 255], drawn from SplitMix64 streams of the image index.  The circuit
    does not depend on the bytes, only the witness does."""
    if image not in ("example", "gradient"):
        raise ValueError("image must be 'example' or 'gradient'")
    ls = log_bytes - 5
    b = 1 << ls
    log_m = (2 * b).bit_length()  # ceil_log2(1 + 2b) for b a power of two
    m = 1 << log_m
    w_cols = min(256, b)
    i = np.arange(b, dtype=np.int64)
    r, c = i // w_cols, i % w_cols
    pred = np.where(r % 2 == 0, np.where(c > 0, i - 1, -1), np.where(r > 0, i - w_cols, -1))
    has = pred >= 0
    one = np.ones(b, np.uint64)
    A = (i.astype(np.uint32), (1 + 2 * i).astype(np.uint32), one)
    B = (i.astype(np.uint32), np.zeros(b, np.uint32), one)
    C = (np.concatenate([i, i[has]]).astype(np.uint32),
         np.concatenate([2 + 2 * i, 1 + 2 * pred[has]]).astype(np.uint32),
         np.ones(b + int(has.sum()), np.uint64))
    wit = np.zeros((32, m), np.uint64)
    nrows = b // w_cols
    with np.errstate(over="ignore"):
        for s in range(32):
            if image == "example":
                seed = (np.uint64(s) << np.uint64(40)) ^ (i.astype(np.uint64) * np.uint64(0x10000001))
                orig = _mix(seed + _G) & np.uint64(0xFF)
            else:
                start = (stream(s, 400, nrows) % np.uint64(256)).astype(np.int64)
                slope = (stream(s, 401, nrows) % np.uint64(9)).astype(np.int64) - 4
                noise = (stream(s, 402, b) % np.uint64(33)).astype(np.int64) - 16
                orig = np.clip(start[r] + slope[r] * c + noise, 0, 255).astype(np.uint64)
            pv = np.where(has, orig[np.maximum(pred, 0)], np.uint64(0))
            filt = (orig - pv) & np.uint64(0xFF)
            wit[s, 0] = 1
            wit[s, 1:1 + 2 * b:2] = orig
            wit[s, 2:2 + 2 * b:2] = filt
    return ls, log_m, A, B, C, wit


def mimd_witnesses(As, Bs, log, k, seed):
    """Fresh satisfying witnesses (inputs from `seed`) for the 32 instance
    circuits (As, Bs) of mimd_instance(log, k, ...)."""
    n = 1 << log
    return np.stack([random_witness(n, k, As[s], Bs[s], seed, tag=500 + s) for s in range(32)])
