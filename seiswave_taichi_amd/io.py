"""Model and data I/O beside the time step: MATLAB models and SEG-Y gathers.

* ``load_mat_model`` follows test3d_homegenous_overhust.py:9-15: the P-wave
  model is the ``Vp`` array of a .mat file, ``vs = vp/1.7`` and Gardner's
  density ``rho = 1.741*(vp/1000)**0.25*1000``; the arrays come back as f32 in
  the file's (nx, ny, nz) order, ready for the ElasticWAVE constructors.
* ``read_segy`` follows ReadSegy.py:3-18 (which needs segyio): traces are
  grouped into shots by their SourceX header word, ``shotnum`` shots of equal
  length, returned as ``data[shot, time, trace]`` with length-1 axes squeezed.
  The reader here is NumPy on the SEG-Y rev 1 layout: 3200-byte text header,
  400-byte big-endian binary header (sample interval at byte 3217, samples per
  trace at 3221, format code at 3225), then per trace a 240-byte header
  (SourceX at byte 73) and the samples; formats 1 (IBM float), 2 (int32),
  3 (int16), 5 (IEEE float) and 8 (int8).
* ``write_segy`` writes the same layout (IBM or IEEE floats), for round trips
  and for handing seismograms to other tools.
"""
from __future__ import annotations

import numpy as np

TEXT_HEADER = 3200
BIN_HEADER = 400
TRACE_HEADER = 240
_FORMATS = {1: (">u4", 4), 2: (">i4", 4), 3: (">i2", 2), 5: (">f4", 4), 8: ("i1", 1)}


def load_mat_model(path, key: str = "Vp", vs_ratio: float = 1.7):
    """(vs, vp, rho) from a MATLAB file holding the P-wave model under ``key``."""
    import scipy.io as sio
    data = sio.loadmat(path)
    if key not in data:
        raise KeyError(f"{path} has no variable {key!r} (variables: {sorted(k for k in data if not k.startswith('__'))})")
    vp = np.asarray(data[key], dtype=np.float64)
    vs = vp / vs_ratio
    rho = 1.741 * (vp / 1000.0) ** 0.25 * 1000.0
    return vs.astype(np.float32), vp.astype(np.float32), rho.astype(np.float32)


def save_mat_model(path, vp, key: str = "Vp"):
    import scipy.io as sio
    sio.savemat(path, {key: np.asarray(vp)})


def ibm_to_ieee(words) -> np.ndarray:
    """IBM System/360 single precision (sign, base-16 exponent biased by 64,
    24-bit fraction) to float32; ``words`` are the big-endian 32-bit words."""
    w = np.asarray(words, dtype=np.uint32)
    sign = np.where((w >> 31) & 1, -1.0, 1.0)
    expo = ((w >> 24) & 0x7F).astype(np.int64) - 64
    frac = (w & 0xFFFFFF).astype(np.float64) / float(1 << 24)
    return (sign * frac * np.power(16.0, expo)).astype(np.float32)


def ieee_to_ibm(values) -> np.ndarray:
    """float values to IBM single-precision words (fraction rounded to nearest)."""
    x = np.asarray(values, dtype=np.float64)
    out = np.zeros(x.shape, dtype=np.uint32)
    nz = x != 0
    a = np.abs(x[nz])
    e = np.floor(np.log2(a) / 4.0).astype(np.int64) + 1          # a = f * 16**e, f in [1/16, 1)
    f = np.rint(a / np.power(16.0, e) * float(1 << 24)).astype(np.int64)
    carry = f >= (1 << 24)                                        # rounding reached 1.0
    f = np.where(carry, f >> 4, f)
    e = np.where(carry, e + 1, e)
    small = f < (1 << 20)                                         # log2 landed one digit high
    f = np.where(small, f << 4, f)
    e = np.where(small, e - 1, e)
    if np.any((e + 64 < 0) | (e + 64 > 127)):
        raise OverflowError("value outside the IBM single-precision range")
    sign = (x[nz] < 0).astype(np.uint32) << 31
    out[nz] = sign | ((e + 64).astype(np.uint32) << 24) | f.astype(np.uint32)
    return out


def read_segy_traces(path):
    """All traces of a SEG-Y file: (samples [ntrace, ns] float32, source_x [ntrace], dt in s)."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < TEXT_HEADER + BIN_HEADER:
        raise ValueError(f"{path}: too short for a SEG-Y file")
    b = raw[TEXT_HEADER:TEXT_HEADER + BIN_HEADER]
    dt_us = int(b[16:18].view(">i2")[0])
    ns = int(b[20:22].view(">i2")[0])
    fmt = int(b[24:26].view(">i2")[0])
    if fmt not in _FORMATS:
        raise ValueError(f"{path}: unsupported SEG-Y sample format {fmt}")
    dtype, size = _FORMATS[fmt]
    rec = TRACE_HEADER + ns * size
    body = raw[TEXT_HEADER + BIN_HEADER:]
    if body.size % rec:
        raise ValueError(f"{path}: {body.size} trace bytes are not a whole number of {rec}-byte traces")
    tr = body.reshape(-1, rec)
    source_x = np.ascontiguousarray(tr[:, 72:76]).view(">i4").ravel().astype(np.int64)
    samples = np.ascontiguousarray(tr[:, TRACE_HEADER:]).view(dtype).reshape(len(tr), ns)
    samples = ibm_to_ieee(samples) if fmt == 1 else samples.astype(np.float32)
    return samples, source_x, dt_us * 1e-6


def read_segy(path, shotnum: int = 0) -> np.ndarray:
    """Shot gathers ``data[shot, time, trace]`` (squeezed), as ReadSegy.read_segy."""
    samples, source_x, _ = read_segy_traces(path)
    trace_num = len(source_x)
    shot_num = shotnum if shotnum else len(set(source_x.tolist()))
    len_shot = trace_num // shot_num
    data = np.zeros((shot_num, samples.shape[1], len_shot))
    for j in range(shot_num):
        data[j] = samples[j * len_shot:(j + 1) * len_shot].T
    return np.squeeze(data)


def write_segy(path, traces, source_x, dt: float, fmt: int = 5) -> None:
    """Write ``traces`` [ntrace, ns] with per-trace SourceX (format 1 IBM or 5 IEEE)."""
    traces = np.asarray(traces, dtype=np.float64)
    ntr, ns = traces.shape
    if fmt not in (1, 5):
        raise ValueError("write_segy writes format 1 (IBM) or 5 (IEEE)")
    text = np.full(TEXT_HEADER, 0x40, dtype=np.uint8)  # EBCDIC blanks
    binh = np.zeros(BIN_HEADER, dtype=np.uint8)
    binh[16:18] = np.frombuffer(np.array([int(round(dt * 1e6))], ">i2").tobytes(), np.uint8)
    binh[20:22] = np.frombuffer(np.array([ns], ">i2").tobytes(), np.uint8)
    binh[24:26] = np.frombuffer(np.array([fmt], ">i2").tobytes(), np.uint8)
    body = np.zeros((ntr, TRACE_HEADER + 4 * ns), dtype=np.uint8)
    hdr = np.zeros((ntr, TRACE_HEADER), dtype=np.uint8)
    hdr[:, 0:4] = np.frombuffer(np.arange(1, ntr + 1, dtype=">i4").tobytes(), np.uint8).reshape(ntr, 4)
    hdr[:, 72:76] = np.frombuffer(np.asarray(source_x, dtype=">i4").tobytes(), np.uint8).reshape(ntr, 4)
    hdr[:, 114:116] = np.frombuffer(np.full(ntr, ns, ">i2").tobytes(), np.uint8).reshape(ntr, 2)
    hdr[:, 116:118] = np.frombuffer(np.full(ntr, int(round(dt * 1e6)), ">i2").tobytes(), np.uint8).reshape(ntr, 2)
    body[:, :TRACE_HEADER] = hdr
    words = ieee_to_ibm(traces).astype(">u4") if fmt == 1 else traces.astype(">f4")
    body[:, TRACE_HEADER:] = np.frombuffer(words.tobytes(), np.uint8).reshape(ntr, 4 * ns)
    with open(path, "wb") as fh:
        fh.write(text.tobytes())
        fh.write(binh.tobytes())
        fh.write(body.tobytes())
