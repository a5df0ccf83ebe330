"""The index-file reader (culsh_amd/csrc/index_file.hpp: what HashTables::load
runs on the host before it touches the device) under AddressSanitizer and
UndefinedBehaviorSanitizer, on a valid file, every truncation boundary,
hostile headers, targeted corruptions and random byte flips.  Every case
must end in a clean status (no sanitizer report, no crash); the targeted
ones must report the expected error."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import oracle
from culsh_amd import workload as wl
from lshtest import write_reference_format

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OK, E_ARG, E_IO, E_RANGE, E_LIMIT = 0, 1, 5, 6, 7


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("san")
    exe = str(d / "index_file_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "culsh_amd", "csrc"),
           os.path.join(ROOT, "tests", "index_file_check.cpp"), "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


@pytest.fixture(scope="module")
def valid_file(tmp_path_factory):
    N, D, L, M, W = 600, 8, 3, 5, 1.0
    X = wl.mixture(N, D, 6, 0.2, 501)
    A, b = wl.projection(L, D, M, W, 502)
    ix = oracle.Index(X, L, M, W, A, b)
    p = tmp_path_factory.mktemp("idx") / "valid.lsh"
    write_reference_format(p, ix, A, b, N, D, L, M, W)
    return p, (N, D, L, M, W), [ix.buckets(t) for t in range(L)]


def run(harness, paths):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    out = subprocess.run([harness, *map(str, paths)], capture_output=True, text=True,
                         timeout=300, env=env)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert "Sanitizer" not in out.stderr and "runtime error" not in out.stderr, out.stderr
    lines = out.stdout.strip().split("\n")
    assert len(lines) == len(paths)
    return [(int(x.split(" ", 1)[0]), x.split(" ", 1)[1] if " " in x else "") for x in lines]


def test_valid_file_reads_clean(harness, valid_file):
    p, _, _ = valid_file
    assert run(harness, [p]) == [(OK, "")]


def test_every_truncation_is_an_io_error(harness, valid_file, tmp_path):
    p, (N, D, L, M, W), B = valid_file
    raw = p.read_bytes()
    cuts = set(range(0, len(raw), 97))
    off = 20
    for t in range(L):  # section boundaries, and one byte either side
        off += 4 * (D * M + M)
        cuts.update((off - 1, off, off + 1))
    off += 4 * L
    cuts.update((off - 1, off, off + 1))
    for t in range(L):
        for n in (N, B[t], B[t], B[t] * M):
            off += 4 * n
            cuts.update((off - 1, off, off + 1))
    paths = []
    for c in sorted(x for x in cuts if 0 <= x < len(raw)):
        q = tmp_path / f"cut{c}.lsh"
        q.write_bytes(raw[:c])
        paths.append(q)
    for rc, msg in run(harness, paths):
        assert rc == E_IO and "truncated" in msg, msg


def test_hostile_headers(harness, valid_file, tmp_path):
    p, (N, D, L, M, W), _ = valid_file
    body = p.read_bytes()[20:]
    cases = {
        "neg_W": (struct.pack("<iiiif", N, D, L, M, -1.0), E_ARG),
        "zero_N": (struct.pack("<iiiif", 0, D, L, M, W), E_ARG),
        "M_65": (struct.pack("<iiiif", N, D, L, 65, W), E_LIMIT),
        "LN_2_31": (struct.pack("<iiiif", 1 << 30, D, 4, M, W), E_LIMIT),
        "huge_D": (struct.pack("<iiiif", N, 1 << 20, L, 64, W), E_IO),  # 268 MB of A claimed
        "huge_N": (struct.pack("<iiiif", 500_000_000, D, 1, M, W), E_IO),
        "nan_W": (struct.pack("<iiii", N, D, L, M) + struct.pack("<I", 0x7FC00000), E_ARG),
    }
    paths, want = [], []
    for name, (hdr, code) in cases.items():
        q = tmp_path / f"{name}.lsh"
        q.write_bytes(hdr + body)
        paths.append(q)
        want.append(code)
    got = run(harness, paths)
    assert [g[0] for g in got] == want, got


def test_targeted_corruptions(harness, valid_file, tmp_path):
    p, (N, D, L, M, W), B = valid_file
    raw = bytearray(p.read_bytes())
    b_off = 20 + 4 * L * (D * M + M)     # B[L]
    t0 = b_off + 4 * L                   # table 0: total[N], start[B0], size[B0], codes
    st0, sz0, cd0 = t0 + 4 * N, t0 + 4 * N + 4 * B[0], t0 + 4 * N + 8 * B[0]

    def patch(off, fmt, val):
        r = bytearray(raw)
        r[off:off + struct.calcsize(fmt)] = struct.pack(fmt, val)
        return bytes(r)

    cases = [
        (patch(b_off, "<I", 0), E_IO, "buckets for N"),
        (patch(b_off, "<I", N + 1), E_IO, "buckets for N"),
        (patch(t0, "<I", N + 5), E_IO, "out of range"),
        (patch(t0, "<I", struct.unpack_from("<I", raw, t0 + 4)[0]), E_IO, "twice"),
        (patch(st0 + 4, "<I", 7), E_IO, "inconsistent"),
        (patch(sz0, "<I", 0), E_IO, "inconsistent"),
        (patch(sz0, "<I", 0xFFFFFFF0), E_IO, "inconsistent"),
        (patch(cd0, "<f", 0.5), E_RANGE, "16-bit"),
        (patch(cd0, "<f", 1e6), E_RANGE, "16-bit"),
        (patch(cd0, "<I", 0x7FC00000), E_RANGE, "16-bit"),
        (bytes(raw) + b"\x00" * 64, OK, ""),  # trailing bytes are ignored, as the reference does
    ]
    paths = []
    for i, (data, _, _) in enumerate(cases):
        q = tmp_path / f"c{i}.lsh"
        q.write_bytes(data)
        paths.append(q)
    for (rc, msg), (_, code, frag) in zip(run(harness, paths), cases):
        assert rc == code and frag in msg, (rc, msg, code, frag)


def test_random_byte_flips_end_cleanly(harness, valid_file, tmp_path):
    p, _, _ = valid_file
    raw = p.read_bytes()
    rng = np.random.default_rng(77)
    paths = []
    for i in range(400):
        r = bytearray(raw)
        for _ in range(int(rng.integers(1, 6))):
            k = int(rng.integers(0, len(r)))
            r[k] = int(rng.integers(0, 256))
        if rng.uniform() < 0.2:
            r = r[:int(rng.integers(0, len(r)))]
        q = tmp_path / f"f{i}.lsh"
        q.write_bytes(bytes(r))
        paths.append(q)
    got = run(harness, paths)
    assert all(rc in (OK, E_ARG, E_IO, E_RANGE, E_LIMIT) for rc, _ in got)
