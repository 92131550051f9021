"""Binary request/response frames for the serving loop (no counterpart in the
reference, which has no server).

All fields are little-endian.  A request frame is a 32-byte header followed
by the series:

    offset  size  field
    0       4     magic b"SSAQ"
    4       2     version (1)
    6       2     sample type (1 = float32; no other type is defined)
    8       8     request id (u64, echoed in the response; it also keys the
                  request's test matrix, so a result does not depend on the
                  wave the request lands in)
    16      4     N, samples in the series (u32, 4 <= N <= MAX_N)
    20      4     L, window length (u32, 2 <= L <= N - 1)
    24      2     k, components (u16, 1 <= k, k + p <= 80 for the server's
                  oversampling p)
    26      2     g, reconstruct components 0..g-1 (u16, 1 <= g <= k)
    28      2     h, forecast horizon (u16)
    30      2     reserved, 0
    32      4 N   the series, float32

A response frame is a 24-byte header followed by the output:

    0       4     magic b"SSAR"
    4       2     version (1)
    6       2     reserved, 0
    8       8     request id
    16      4     status (i32): 0 served, -1 the decomposition or forecast
                  failed, -3 refused (larger than the device budget),
                  -4 rejected (a field outside the ranges above, or a horizon
                  other than the server's)
    20      4     n_out (u32): N + h when served, else 0
    24      8 n_out  reconstruction of the group (N values) then the h-step
                  forecast, float64

A frame whose magic, version or sample type is wrong cannot be resynchronised:
`read_request` raises `WireError` and the server closes the stream.  A frame
that parses but whose field values are out of range is answered with status
-4 and the stream goes on.
"""
import struct

import numpy as np

MAGIC_REQ = b"SSAQ"
MAGIC_RESP = b"SSAR"
VERSION = 1
SAMPLE_F32 = 1
MAX_N = 1 << 24
MAX_BLOCK = 80      # k + p, the device's widest block

REQ_HEADER = struct.Struct("<4sHHQIIHHHH")
RESP_HEADER = struct.Struct("<4sHHQiI")
assert REQ_HEADER.size == 32 and RESP_HEADER.size == 24

STATUS_OK, STATUS_FAILED, STATUS_REFUSED, STATUS_REJECTED = 0, -1, -3, -4


class WireError(ValueError):
    """A frame that cannot be parsed (the stream cannot continue)."""


def encode_request(rid, x, L, k, g, h):
    x = np.ascontiguousarray(x, dtype="<f4")
    return REQ_HEADER.pack(MAGIC_REQ, VERSION, SAMPLE_F32, int(rid), len(x), int(L), int(k), int(g), int(h), 0) + \
        x.tobytes()


def check_fields(N, L, k, g, h, p=10):
    """None if the fields are in range for oversampling p, else the reason."""
    if not 4 <= N <= MAX_N:
        return f"N = {N} outside [4, {MAX_N}]"
    if not 2 <= L <= N - 1:
        return f"L = {L} outside [2, N - 1]"
    if not 1 <= k <= MAX_BLOCK - p:
        return f"k = {k} outside [1, {MAX_BLOCK - p}]"
    if not 1 <= g <= k:
        return f"g = {g} outside [1, k]"
    if h < 0:
        return f"h = {h} negative"
    return None


def _read_exact(stream, n):
    """n bytes from a blocking binary stream; b"" at a clean end of stream
    (no byte of the frame read), WireError if it ends inside the frame."""
    buf = bytearray()
    while len(buf) < n:
        chunk = stream.read(n - len(buf))
        if not chunk:
            if not buf:
                return b""
            raise WireError(f"stream ended inside a frame ({len(buf)} of {n} bytes)")
        buf += chunk
    return bytes(buf)


def parse_request_header(hdr):
    """(rid, N, L, k, g, h) of a 32-byte request header; WireError if its
    magic, version or sample type is wrong."""
    magic, ver, stype, rid, N, L, k, g, h, _ = REQ_HEADER.unpack(hdr)
    if magic != MAGIC_REQ:
        raise WireError(f"bad request magic {magic!r}")
    if ver != VERSION:
        raise WireError(f"unsupported version {ver}")
    if stype != SAMPLE_F32:
        raise WireError(f"unsupported sample type {stype}")
    if N > MAX_N:
        raise WireError(f"N = {N} above {MAX_N}")
    return rid, N, L, k, g, h


def read_request(stream):
    """The next request frame as (rid, x float32[N], L, k, g, h), or None at a
    clean end of stream."""
    hdr = _read_exact(stream, REQ_HEADER.size)
    if not hdr:
        return None
    rid, N, L, k, g, h = parse_request_header(hdr)
    body = _read_exact(stream, 4 * N)
    if len(body) != 4 * N:
        raise WireError("stream ended before the series")
    return rid, np.frombuffer(body, dtype="<f4").copy(), L, k, g, h


def encode_response(rid, status, out=None):
    out = np.zeros(0) if out is None or status != STATUS_OK else np.ascontiguousarray(out, dtype="<f8")
    return RESP_HEADER.pack(MAGIC_RESP, VERSION, 0, int(rid), int(status), len(out)) + out.tobytes()


def read_response(stream):
    """(rid, status, out float64[n_out]) or None at a clean end of stream."""
    hdr = _read_exact(stream, RESP_HEADER.size)
    if not hdr:
        return None
    magic, ver, _, rid, status, n = RESP_HEADER.unpack(hdr)
    if magic != MAGIC_RESP:
        raise WireError(f"bad response magic {magic!r}")
    if ver != VERSION:
        raise WireError(f"unsupported version {ver}")
    body = _read_exact(stream, 8 * n)
    if len(body) != 8 * n:
        raise WireError("stream ended before the output")
    return rid, status, np.frombuffer(body, dtype="<f8").copy()
