"""guide_ref.py - CPU ORACLE for the learned directional proposal (guided form).

TEST INFRASTRUCTURE ONLY: the checker for the HIP MLP kernels (forward, loss, backward, Adam) and
for the von Mises-Fisher mixture sampling/pdf.  Only tests/ may import it; the product library
never calls it.

The reference renderer has no learned proposal, so this file restates the specification the HIP
kernels implement (there is no reference file:line to follow):

  inputs   x in [0,1]^9 (position in scene bounds, outgoing direction and shading normal mapped
           from [-1,1]), clamped to [0,1]
  encoding e = [x (9), for input i, frequency k<16: sin(2 pi frac(x_i f_k)), cos(...)] with
           f_k = 2^(k/2) in fp32 (half-octave steps, 1 .. 181) -> 297 features, zero padded to 304
  network  h0 = bf16(e); z_l = h_{l-1} W_l + b_l (fp32 accumulate); h_l = bf16(relu(z_l)), l=1..4;
           o = h4 W5 + b5 (fp32), 48 outputs of which 40 are used; W_l in bf16, b_l in fp32
  mixture  8 lobes, lobe j reads o[5j .. 5j+4] = (logit, r_xyz, s): pi = softmax(logits),
           mu = r/|r|, kappa = exp(clamp(s, S_MIN, S_MAX));
           vMF(w) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1))
  loss     L = (1/B) sum_m -wt_m log q(w_m | x_m)  (the KL objective of neural importance sampling
           estimated with the tracer's own samples: wt = luminance(f Li cos / p_sample))
  backward d5 = bf16(dL/do); dh_l = d_{l+1} W_{l+1}^T (fp32); d_l = bf16(dh_l * [h_l > 0]);
           dW_l = sum_m h_{l-1}^T d_l, db_l = sum_m d_l (fp32)
  Adam     m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; w -= lr (m/(1-b1^t)) / (sqrt(v/(1-b2^t)) + eps)

Rounding to bf16 is round-to-nearest-even on the fp32 value, modelled exactly here; the rest runs
in float64, so the HIP results differ only by fp32 accumulation order and device sin/cos.

Pinning: the backward pass is checked against central finite differences of the forward/loss in
float64 without bf16 rounding; the vMF pdf integrates to 1 (quadrature), its mean direction has
the closed-form length coth(kappa) - 1/kappa, and the sampling routine's histogram matches the pdf.
"""
import numpy as np

N_IN, N_FREQ = 9, 16
ENC = N_IN + N_IN * N_FREQ * 2          # 297
K1 = 304
H = 64
N_LOBES = 8
N_OUT = 5 * N_LOBES                     # 40
N_OUT_PAD = 48
S_MIN, S_MAX = -3.0, 7.0

# canonical parameter order: W1 [K1][H], b1, W2 [H][H], b2, W3, b3, W4, b4, W5 [H][N_OUT_PAD], b5
SHAPES = [(K1, H), (H,), (H, H), (H,), (H, H), (H,), (H, H), (H,), (H, N_OUT_PAD), (N_OUT_PAD,)]
PARAM_COUNT = int(sum(np.prod(s) for s in SHAPES))   # 35120


def unpack(flat):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(flat[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def pack(arrs):
    return np.concatenate([np.asarray(a, dtype=np.float64).reshape(-1) for a in arrs])


def bf16(x):
    """Round-to-nearest-even of float32(x) to bfloat16, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    r = u.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(nan, np.nan, r)


def freq(k):
    """f_k = 2^(k/2) as the fp32 value the kernels use."""
    base = np.float32(np.sqrt(2.0)) if k & 1 else np.float32(1.0)
    return np.float32(base * np.float32(2.0 ** (k >> 1)))


def encode(x):
    x = np.clip(np.asarray(x, dtype=np.float32), 0.0, 1.0)
    n = x.shape[0]
    e = np.zeros((n, K1), dtype=np.float64)
    e[:, :N_IN] = x
    for i in range(N_IN):
        for k in range(N_FREQ):
            t = x[:, i].astype(np.float32) * freq(k)                     # one fp32 rounding
            f = (t - np.floor(t)).astype(np.float64)
            e[:, N_IN + 32 * i + 2 * k] = np.sin(2 * np.pi * f)
            e[:, N_IN + 32 * i + 2 * k + 1] = np.cos(2 * np.pi * f)
    return e


def forward(params, x, rounding=True):
    """Returns (o [n][48], acts) with acts = [h0..h4] as used by the next layer."""
    W1, b1, W2, b2, W3, b3, W4, b4, W5, b5 = unpack(params)
    rd = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    Ws = [rd(W1), rd(W2), rd(W3), rd(W4), rd(W5)]
    bs = [b1, b2, b3, b4, b5]
    h = rd(encode(x))
    acts = [h]
    for l in range(4):
        z = h @ Ws[l] + bs[l]
        h = rd(np.maximum(z, 0.0))
        acts.append(h)
    o = h @ Ws[4] + bs[4]
    return o, acts


def mixture(o):
    o = np.asarray(o, dtype=np.float64)[:, :N_OUT].reshape(-1, N_LOBES, 5)
    logits = o[:, :, 0]
    r = o[:, :, 1:4]
    s = o[:, :, 4]
    lm = logits - logits.max(axis=1, keepdims=True)
    pi = np.exp(lm) / np.exp(lm).sum(axis=1, keepdims=True)
    rn = np.sqrt((r * r).sum(axis=2, keepdims=True))
    mu = r / np.maximum(rn, 1e-12)
    kappa = np.exp(np.clip(s, S_MIN, S_MAX))
    return pi, mu, kappa, rn[:, :, 0], s


def log_vmf(mu, kappa, w):
    """log vMF density, mu [n][L][3], kappa [n][L], w [n][3] -> [n][L]."""
    c = (mu * w[:, None, :]).sum(axis=2)
    return np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa)) + kappa * (c - 1.0)


def mixture_pdf(o, w):
    pi, mu, kappa, _, _ = mixture(o)
    lv = log_vmf(mu, kappa, np.asarray(w, dtype=np.float64))
    return (pi * np.exp(lv)).sum(axis=1)


def loss_and_dout(o, w, wt, batch=None):
    """Mean loss and dL/do [n][48] for a batch (batch size B defaults to n)."""
    n = o.shape[0]
    B = n if batch is None else batch
    w = np.asarray(w, dtype=np.float64)
    wt = np.asarray(wt, dtype=np.float64)
    pi, mu, kappa, rn, s = mixture(o)
    lv = log_vmf(mu, kappa, w)
    a = np.log(pi) + lv
    amax = a.max(axis=1, keepdims=True)
    logq = amax[:, 0] + np.log(np.exp(a - amax).sum(axis=1))
    gam = np.exp(a - logq[:, None])
    g = -wt / B
    c = (mu * w[:, None, :]).sum(axis=2)
    d = np.zeros((n, N_LOBES, 5))
    d[:, :, 0] = g[:, None] * (gam - pi)
    perp = w[:, None, :] - c[:, :, None] * mu
    d[:, :, 1:4] = (g[:, None] * gam * kappa / np.maximum(rn, 1e-12))[:, :, None] * perp
    dlk = 1.0 - 2.0 * kappa / np.expm1(2.0 * kappa) + kappa * (c - 1.0)
    inside = (s > S_MIN) & (s < S_MAX)
    d[:, :, 4] = g[:, None] * gam * dlk * inside
    dout = np.zeros((n, N_OUT_PAD))
    dout[:, :N_OUT] = d.reshape(n, N_OUT)
    return float((-wt * logq).sum() / B), dout


def gradients(params, x, w, wt, rounding=True):
    """(loss, flat gradient in canonical order) for one batch."""
    W1, b1, W2, b2, W3, b3, W4, b4, W5, b5 = unpack(params)
    rd = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    Ws = [rd(W1), rd(W2), rd(W3), rd(W4), rd(W5)]
    o, acts = forward(params, x, rounding)
    loss, dout = loss_and_dout(o, w, wt)
    d = rd(dout)
    gW = [None] * 5
    gb = [None] * 5
    for l in range(4, -1, -1):
        gW[l] = acts[l].T @ d
        gb[l] = d.sum(axis=0)
        if l == 0:
            break
        dh = d @ Ws[l].T
        d = rd(dh * (acts[l] > 0))
    gW[0][ENC:, :] = 0.0          # padding rows carry no parameters
    gW[4][:, N_OUT:] = 0.0
    gb[4][N_OUT:] = 0.0
    return loss, pack([gW[0], gb[0], gW[1], gb[1], gW[2], gb[2], gW[3], gb[3], gW[4], gb[4]])


def loss_only(params, x, w, wt, rounding=False):
    o, _ = forward(params, x, rounding)
    return loss_and_dout(o, w, wt)[0]


def adam(params, grads, m, v, t, lr, b1=0.9, b2=0.99, eps=1e-8):
    """One Adam step (t = 1-based step count); returns new (params, m, v)."""
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def pad_mask():
    """1 for real parameters, 0 for the padding entries of W1 rows >= 297 and W5/b5 columns >= 40."""
    W1 = np.ones((K1, H)); W1[ENC:, :] = 0
    W5 = np.ones((H, N_OUT_PAD)); W5[:, N_OUT:] = 0
    b5 = np.ones(N_OUT_PAD); b5[N_OUT:] = 0
    return pack([W1, np.ones(H), np.ones((H, H)), np.ones(H), np.ones((H, H)), np.ones(H),
                 np.ones((H, H)), np.ones(H), W5, b5])


def sample_vmf(mu, kappa, u1, u2):
    """Inverse-CDF vMF sample around unit mu (float64, one lobe per row)."""
    wz = 1.0 + np.log1p(u1 * np.expm1(-2.0 * kappa)) / kappa
    wz = np.clip(wz, -1.0, 1.0)
    r = np.sqrt(np.maximum(0.0, 1.0 - wz * wz))
    phi = 2 * np.pi * u2
    t, b = frame(mu)
    return wz[:, None] * mu + (r * np.cos(phi))[:, None] * t + (r * np.sin(phi))[:, None] * b


def frame(n):
    """Orthonormal (t, b) around unit n (branchless construction of Duff et al. 2017)."""
    s = np.where(n[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + n[:, 2])
    bb = n[:, 0] * n[:, 1] * a
    t = np.stack([1.0 + s * n[:, 0] * n[:, 0] * a, s * bb, -s * n[:, 0]], axis=1)
    b = np.stack([bb, s + n[:, 1] * n[:, 1] * a, -n[:, 1]], axis=1)
    return t, b
