"""CPU oracle of the ionosphere warm-start network (test infrastructure only).

The network has no counterpart in the reference: it is the learned initial
guess for the polar potential solve (NO/nonorthogonal_mi.c:449-464 is the
reference's own guess, the previous potential). This file states, in numpy,
exactly what the device kernels compute, so that tests can compare the GPU
forward pass, gradients and optimizer step with it. It is pinned by finite
differences in float64 (tests/test_unet_oracle.py) and by direct-loop
convolutions on small inputs.

Network (a 4-level U-Net, 32 base channels), NHWC tensors, H = colatitude
rows, W = longitude columns (periodic), N = samples:

    x    [N,H,W,32]  input features (bf16)
    e0a  = conv(x, 32->32);        e0 = conv(e0a, 32->32)
    e1a  = conv(pool(e0), 32->64); e1 = conv(e1a, 64->64)
    e2a  = conv(pool(e1), 64->128); e2 = conv(e2a, 128->128)
    e3a  = conv(pool(e2), 128->256); e3 = conv(e3a, 256->256)
    d2a  = conv(cat(up(e3), e2), 384->128); d2 = conv(d2a, 128->128)
    d1a  = conv(cat(up(d2), e1), 192->64);  d1 = conv(d1a, 64->64)
    d0a  = conv(cat(up(d1), e0), 96->32);   d0 = conv(d0a, 32->32)
    out  = head(d0): 1x1, 32->1, float32

conv = 3x3, stride 1, zero padding in colatitude, periodic in longitude,
bias, ReLU, output rounded to bf16. Weights of a conv are [cout][3][3][cin]
float32 masters, used as bf16. pool = 2x2 mean (bf16 out), up = nearest.
Products are bf16 x bf16 accumulated in float32.

Backward: gz = bf16(g * (out > 0)); bias grad = sum gz; weight grad =
sum over pixels of gz * input patch; input grad = conv of gz with the
flipped, channel-transposed weights (float32). The loss is the mean square
error of `out` against the target over all pixels.
"""
import numpy as np

LEVELS = (32, 64, 128, 256)
IN_CHANNELS = 32

# (name, cin, cout, input): input is a tensor name or ("cat", coarse, skip)
# or ("pool", name).
LAYERS = (
    ("e0a", IN_CHANNELS, 32, "x"),
    ("e0", 32, 32, "e0a"),
    ("e1a", 32, 64, ("pool", "e0")),
    ("e1", 64, 64, "e1a"),
    ("e2a", 64, 128, ("pool", "e1")),
    ("e2", 128, 128, "e2a"),
    ("e3a", 128, 256, ("pool", "e2")),
    ("e3", 256, 256, "e3a"),
    ("d2a", 384, 128, ("cat", "e3", "e2")),
    ("d2", 128, 128, "d2a"),
    ("d1a", 192, 64, ("cat", "d2", "e1")),
    ("d1", 64, 64, "d1a"),
    ("d0a", 96, 32, ("cat", "d1", "e0")),
    ("d0", 32, 32, "d0a"),
)
HEAD_IN = 32


def bf16(a):
    """Round float32 values to the nearest bf16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = rounded.astype(np.uint32).view(np.float32).copy()
    nan = np.isnan(a)
    out[nan] = np.float32("nan")
    return out


class Identity:
    """No rounding: the float64 form used by finite-difference checks."""

    @staticmethod
    def round(a):
        return a


class Bf16:
    @staticmethod
    def round(a):
        return bf16(a)


def patches(x, up=None):
    """[N,H,W,9*C] 3x3 patches of x (zero rows outside, periodic columns),
    k index = (ky*3 + kx)*C + c."""
    n, h, w, c = x.shape
    pad = np.zeros((n, h + 2, w, c), dtype=x.dtype)
    pad[:, 1:-1] = x
    cols = []
    for ky in range(3):
        rows = pad[:, ky:ky + h]
        for kx in range(3):
            cols.append(np.roll(rows, 1 - kx, axis=2))
    return np.concatenate(cols, axis=3)


def upsample(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def pool(x):
    n, h, w, c = x.shape
    return 0.25 * x.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def conv_forward(x, wt, b, fmt):
    """x [N,H,W,cin], wt [cout,3,3,cin] -> relu(conv + b), rounded."""
    cout = wt.shape[0]
    p = patches(x)
    acc = p.reshape(-1, p.shape[-1]) @ wt.reshape(cout, -1).T + b
    acc = acc.reshape(x.shape[:3] + (cout,))
    return fmt.round(np.maximum(acc, 0.0).astype(x.dtype))


def conv_input_grad(gz, wt):
    """Adjoint of the patch product: conv of gz with flipped, transposed wt."""
    flipped = wt[:, ::-1, ::-1, :].transpose(3, 1, 2, 0)  # [cin,3,3,cout]
    cin = flipped.shape[0]
    p = patches(gz)
    return (p.reshape(-1, p.shape[-1]) @ flipped.reshape(cin, -1).T).reshape(
        gz.shape[:3] + (cin,))


def conv_weight_grad(x, gz):
    p = patches(x)
    cout = gz.shape[-1]
    return (gz.reshape(-1, cout).T @ p.reshape(-1, p.shape[-1])).reshape(
        cout, 3, 3, x.shape[-1])


def init_params(seed):
    """He-normal conv weights, zero biases, small head (same draws as the
    device initialiser given the same float32 arrays)."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, cin, cout, _ in LAYERS:
        std = np.sqrt(2.0 / (9.0 * cin))
        params[name + ".w"] = (rng.standard_normal((cout, 3, 3, cin)) * std).astype(np.float32)
        params[name + ".b"] = np.zeros(cout, dtype=np.float32)
    params["head.w"] = (rng.standard_normal(HEAD_IN) * 0.01).astype(np.float32)
    params["head.b"] = np.zeros(1, dtype=np.float32)
    return params


def layer_input(acts, spec, fmt):
    if isinstance(spec, str):
        return acts[spec]
    if spec[0] == "pool":
        return fmt.round(pool(acts[spec[1]]))
    return np.concatenate([upsample(acts[spec[1]]), acts[spec[2]]], axis=3)


def forward(params, x, fmt=Bf16, dtype=np.float32):
    """Returns (out [N,H,W], activations)."""
    acts = {"x": fmt.round(np.asarray(x, dtype=dtype))}
    for name, _, _, spec in LAYERS:
        inp = layer_input(acts, spec, fmt)
        w = fmt.round(params[name + ".w"].astype(dtype))
        acts[name] = conv_forward(inp, w, params[name + ".b"].astype(dtype), fmt)
    out = acts["d0"] @ params["head.w"].astype(dtype) + params["head.b"].astype(dtype)[0]
    return out, acts


def loss_and_grad_out(out, target):
    diff = out - target
    return float(np.mean(diff.astype(np.float64) ** 2)), (2.0 / diff.size) * diff


def backward(params, acts, g_out, fmt=Bf16, dtype=np.float32):
    """Gradients of every parameter for the upstream gradient g_out."""
    grads = {}
    d0 = acts["d0"]
    grads["head.w"] = np.einsum("nhw,nhwc->c", g_out, d0).astype(dtype)
    grads["head.b"] = np.array([g_out.sum()], dtype=dtype)
    g = {name: np.zeros(acts[name].shape, dtype=dtype) for name in acts}
    g["d0"] += g_out[..., None] * params["head.w"].astype(dtype)
    for name, cin, cout, spec in reversed(LAYERS):
        gz = fmt.round((g[name] * (acts[name] > 0)).astype(dtype))
        inp = layer_input(acts, spec, fmt)
        grads[name + ".b"] = gz.reshape(-1, cout).sum(axis=0).astype(dtype)
        grads[name + ".w"] = conv_weight_grad(inp, gz).astype(dtype)
        if spec == "x":
            continue
        w = fmt.round(params[name + ".w"].astype(dtype))
        gin = conv_input_grad(gz, w).astype(dtype)
        if isinstance(spec, str):
            g[spec] += gin
        elif spec[0] == "pool":
            g[spec[1]] += 0.25 * upsample(gin)
        else:
            coarse, skip = spec[1], spec[2]
            cc = acts[coarse].shape[-1]
            gc = gin[..., :cc]
            n, h, w_, _ = gc.shape
            g[coarse] += gc.reshape(n, h // 2, 2, w_ // 2, 2, cc).sum(axis=(2, 4))
            g[skip] += gin[..., cc:]
    return grads


def adam_step(params, grads, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (bias-corrected), float32, in place."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    c1 = 1.0 - beta1 ** t
    c2 = 1.0 - beta2 ** t
    for k, g in grads.items():
        m = state.setdefault("m." + k, np.zeros_like(params[k]))
        v = state.setdefault("v." + k, np.zeros_like(params[k]))
        m[...] = (beta1 * m + (1.0 - beta1) * g).astype(np.float32)
        v[...] = (beta2 * v + (1.0 - beta2) * g * g).astype(np.float32)
        step = lr * (m / c1) / (np.sqrt(v / c2) + eps)
        params[k][...] = (params[k] - step).astype(np.float32)
