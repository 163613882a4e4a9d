"""Python mirror of the denoiser entry points of include/bte_api.h.

The U-Net fills the reference's denoiser slot (src/host/denoiser.cu:72-116:
cudaToggleDenoiser, denoiseAndPresent).  Every call goes through libbte.so;
there is no host fallback.
"""
import ctypes
import os

import numpy as np

from .renderer import _call, _ptr, lib

WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_c4.npy")
OUTPUT_TENSOR = 39
# (name, input channels, output channels, level) of the 15 convolutions
LAYERS = [("c0a", 9, 32, 0), ("c0b", 32, 32, 0), ("c1a", 32, 64, 1), ("c1b", 64, 64, 1), ("c2a", 64, 128, 2),
          ("c2b", 128, 128, 2), ("c3a", 128, 128, 3), ("c3b", 128, 128, 3), ("d2a", 256, 128, 2),
          ("d2b", 128, 128, 2), ("d1a", 192, 64, 1), ("d1b", 64, 64, 1), ("d0a", 96, 32, 0), ("d0b", 32, 32, 0),
          ("out", 32, 3, 0)]


def flops_per_pixel():
    """Multiply-adds x 2 of one forward pass per full-resolution pixel."""
    return sum(2.0 * 9 * ci * co / 4 ** lvl for _, ci, co, lvl in LAYERS)


def param_count():
    return int(lib().bte_denoiser_param_count())


def init(seed=0):
    _call("bte_denoiser_init", int(seed))


def release():
    _call("bte_denoiser_release")


def set_weights(flat):
    w = np.ascontiguousarray(flat, np.float32)
    _call("bte_denoiser_set_weights", _ptr(w), w.size)


def get_weights():
    out = np.empty(param_count(), np.float32)
    _call("bte_denoiser_get_weights", _ptr(out), out.size)
    return out


def get_grads():
    out = np.empty(param_count(), np.float32)
    _call("bte_denoiser_get_grads", _ptr(out), out.size)
    return out


def load(path=WEIGHTS_PATH):
    """Loads trained weights (a float32 .npy in canonical order) through the library's own reader."""
    _call("bte_denoiser_load_file", os.fsencode(path))


def save(path=WEIGHTS_PATH):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, get_weights(), allow_pickle=False)


def forward(x9):
    """x9 (n, h, w, 9) float32 -> network output (n, h, w, 3), log(1 + RGB)."""
    x = np.ascontiguousarray(x9, np.float32)
    n, h, w, _ = x.shape
    out = np.empty((n, h, w, 3), np.float32)
    _call("bte_denoiser_forward_host", _ptr(x), n, h, w, _ptr(out))
    return out


def train_step(x9, target3, lr=0.0, weight_decay=0.0):
    x = np.ascontiguousarray(x9, np.float32)
    t = np.ascontiguousarray(target3, np.float32)
    n, h, w, _ = x.shape
    loss = np.zeros(1, np.float32)
    _call("bte_denoiser_train_step_host", _ptr(x), _ptr(t), n, h, w, float(lr), float(weight_decay), _ptr(loss))
    return float(loss[0])


def tensor(tid):
    shape = np.zeros(4, np.int32)
    _call("bte_denoiser_tensor_shape", int(tid), _ptr(shape))
    out = np.empty(tuple(int(s) for s in shape), np.float32)
    _call("bte_denoiser_read_tensor", int(tid), _ptr(out), out.size)
    return out


LOSS_L1_LOG, LOSS_RELMSE = 0, 1


def set_loss(mode):
    _call("bte_denoiser_set_loss", int(mode))


def set_flips(flips):
    """Inference flip ensemble: 1, 2 or 4 mirrored copies averaged in log space."""
    _call("bte_denoiser_set_flips", int(flips))


def set_conv_kernel(variant):
    """0: fragments from L1/L2, 1: LDS-staged weight tiles, 2 (default): LDS-staged input tiles."""
    _call("bte_denoiser_set_conv_kernel", int(variant))


def get_flips():
    return _call("bte_denoiser_get_flips")


def _hip():
    """The HIP runtime libbte is linked against (already loaded with it), for
    the device buffers of denoise_image."""
    lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def denoise_image(rgb, albedo, normal):
    """Denoises host images (H, W, 3 each, row 0 = bottom as Renderer.accum)
    through bte_denoise_image (device float4 buffers)."""
    H, W, _ = np.shape(rgb)
    hip = _hip()
    nbytes = H * W * 16
    bufs = []
    try:
        for a in (rgb, albedo, normal, None):
            p = ctypes.c_void_p()
            if hip.hipMalloc(ctypes.byref(p), nbytes) != 0:
                raise MemoryError("hipMalloc failed")
            bufs.append(p)
            if a is not None:
                t = np.zeros((H, W, 4), np.float32)
                t[..., :3] = a
                if hip.hipMemcpy(p, t.ctypes.data_as(ctypes.c_void_p), nbytes, 1) != 0:
                    raise RuntimeError("hipMemcpy to the device failed")
        _call("bte_denoise_image", bufs[0], bufs[1], bufs[2], W, H, bufs[3])
        out = np.empty((H, W, 4), np.float32)
        if hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), bufs[3], nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy from the device failed")
        return out[..., :3].copy()
    finally:
        for p in bufs:
            hip.hipFree(p)


def dataset_reserve(slots, width, height):
    _call("bte_denoiser_dataset_reserve", int(slots), int(width), int(height))


def capture(slot, which):
    _call("bte_denoiser_capture", int(slot), int(which))


def train(steps, batch, lr=1e-3, weight_decay=1e-4, seed=0):
    losses = np.zeros(int(steps), np.float32)
    _call("bte_denoiser_train", int(steps), int(batch), float(lr), float(weight_decay), ctypes.c_uint(seed),
          _ptr(losses))
    return losses


def denoise():
    _call("bte_denoise")


def toggle():
    return _call("bte_toggle_denoiser") == 1


def denoised(width, height):
    out = np.empty((height, width, 3), np.float32)
    _call("bte_read_denoised", _ptr(out), out.size)
    return out
