"""Test infrastructure only: CPU restatement of the reference's training-record
encoders (scripts/summit_scripts/utils.py:60-118) and of its LMDB header
block (scripts/summit_scripts/sim_prod.py:186-197). Used to check
namsa_amd.records; nothing in the product imports this module.

The reference computes with numpy on float32 arrays and calls
scipy.ndimage.gaussian_filter (both present here). The restatement keeps
that dtype flow.

Pins (tests/test_records.py):
- process_cbed: a hand-worked item ([0, 1, 2, 3] -> +-1.3416, +-0.4472);
  mean 0 and deviation 1 per item; invariance under x -> a x + b
  (a > 0); a constant item gives NaN.
- process_potential: gaussian_filter against an explicit convolution with
  'symmetric' padding and the normalised exp(-k^2 / 2 sigma^2) weights;
  the 128 x 128 record at 256^2 that sim_prod.py:127 checks for; a
  constant field gives zeros (deviation floored at 1/sqrt(size)).
- lmdb_headers / record keys: byte layout worked out from the source
  lines (int64 shapes, 'float16' names, the six zero bytes that
  bytes(len(headers)) makes).
"""
import numpy as np
from scipy.ndimage import gaussian_filter


def process_cbed(cbed):
    """utils.py:104-118: per item (axes 1 and -1) standardisation, float16."""
    cbed = np.asarray(cbed, dtype=np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        out = (cbed - np.mean(cbed, axis=(1, -1), keepdims=True)) / np.std(cbed, axis=(1, -1), keepdims=True)
    return out.astype(np.float16)


def crop_slice(shape):
    """utils.py:77: slice(shape[0] // 4, 3 * shape[1] // 4), used on both axes."""
    return slice(int(shape[0] // 4), int(3 * shape[1] // 4))


def process_potential(pot_slices, sigma=1.2):
    """utils.py:74-102: imag mean over slices, Gaussian filter, crop,
    standardise with the deviation floored at 1/sqrt(size), float16, [1][rows][cols]."""
    proj = np.imag(np.asarray(pot_slices, dtype=np.complex64)).mean(0)
    proj = gaussian_filter(proj, sigma)
    s = crop_slice(proj.shape)
    proj = proj[s, s]
    proj = (proj - proj.mean()) / max(proj.std(), 1.0 / np.sqrt(proj.size))
    return np.expand_dims(proj, axis=0).astype(np.float16)


def gaussian_direct(img, sigma=1.2, truncate=4.0):
    """Separable Gaussian by explicit convolution: weights exp(-k^2 / 2 sigma^2)
    normalised over |k| <= int(truncate sigma + 0.5), half-sample symmetric
    edges (numpy 'symmetric' padding), axis 0 then axis 1, float32 between
    the passes (as scipy stores into a float32 output)."""
    r = int(truncate * sigma + 0.5)
    k = np.arange(-r, r + 1)
    w = np.exp(-0.5 * k * k / sigma ** 2)
    w = w / w.sum()
    out = np.asarray(img, dtype=np.float32)
    for axis in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[axis] = (r, r)
        p = np.pad(out.astype(np.float64), pad, mode="symmetric")
        n = out.shape[axis]
        acc = sum(w[j] * np.take(p, np.arange(j, j + n), axis=axis) for j in range(2 * r + 1))
        out = acc.astype(np.float32)
    return out


def lmdb_headers(input_shape=(64, 256, 256), output_shape=(1, 128, 128), record_names=("2d_potential_", "cbed_")):
    """sim_prod.py:190-197: header entries as (key, value) bytes, in order,
    then b'header_entries' -> bytes(6) (six zero bytes)."""
    h = [(b"input_dtype", b"float16"),
         (b"input_shape", np.array(input_shape, dtype=np.int64).tobytes()),
         (b"output_shape", np.array(output_shape, dtype=np.int64).tobytes()),
         (b"output_dtype", b"float16"),
         (b"output_name", record_names[0].encode("ascii")),
         (b"input_name", record_names[1].encode("ascii"))]
    return h + [(b"header_entries", b"\x00" * len(h))]
