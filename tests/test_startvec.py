"""Start-vector files of the C1 campaign (host logic of the HIP library, no
GPU): the 17-significant-digit text format and its reader
(materialize_coarse_start.cpp:18-43, run_campaign.cpp:287-302), SHA-256
against FIPS 180-4 test vectors and Python's hashlib, and the seeded
perturbation (run_campaign.cpp:304-327) against a restatement of the C++
standard library's mt19937_64 + normal_distribution (Marsaglia polar method
with the saved second value, generate_canonical = u64 / 2^64)."""
import hashlib
import math
import os

import numpy as np
import pytest

import numa_hip as H


def test_sha256_known_answers():
    # FIPS 180-4 / NIST example messages
    assert H.sha256_bytes(b"") == \
        "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"
    assert H.sha256_bytes(b"abc") == \
        "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    assert H.sha256_bytes(b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq") == \
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"


@pytest.mark.parametrize("size", [1, 55, 56, 63, 64, 65, 119, 120, 1000, 70001])
def test_sha256_matches_hashlib(tmp_path, size):
    data = np.random.default_rng(size).integers(0, 256, size, dtype=np.uint8).tobytes()
    assert H.sha256_bytes(data) == hashlib.sha256(data).hexdigest()
    p = tmp_path / "blob.bin"
    p.write_bytes(data)
    assert H.sha256_file(str(p)) == hashlib.sha256(data).hexdigest()


def test_start_file_round_trip_and_format(tmp_path):
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.normal(size=500) * 10.0 ** rng.integers(-300, 300, 500),
                        [0.0, -0.0, 5e-324, -1.7976931348623157e308, 1.0 / 3.0, 9.0522684695809268]])
    p = str(tmp_path / "start.txt")
    H.write_start(p, x)
    text = open(p).read()
    assert text == "".join("%.17g\n" % v for v in x)
    y = H.read_start(p, x.size)
    assert np.array_equal(y.view(np.uint64), x.view(np.uint64))  # bit for bit, signed zero too
    assert H.sha256_file(p) == hashlib.sha256(text.encode()).hexdigest()


def test_start_reader_accepts_whitespace_and_rejects_bad_files(tmp_path):
    p = tmp_path / "s.txt"
    p.write_text("1.5  2\t-3e-2\n\n4\n")
    assert list(H.read_start(str(p), 4)) == [1.5, 2.0, -3e-2, 4.0]
    for bad, n in (("1\n2\n", 3), ("1\n2\n3\n", 2), ("1\nnan\n", 2), ("1\ninf\n", 2),
                   ("1\nabc\n", 2), ("1\n2x\n", 2)):
        p.write_text(bad)
        with pytest.raises(H.NumaError):
            H.read_start(str(p), n)
    with pytest.raises(H.NumaError):
        H.read_start(str(tmp_path / "missing.txt"), 1)


def _mt19937_64(seed):
    """std::mt19937_64 (the 64-bit Mersenne twister of [rand.predef])."""
    w, n, m, r = 64, 312, 156, 31
    a, u, d, s, b, t, c, l, f = (0xB5026F5AA96619E9, 29, 0x5555555555555555, 17,
                                 0x71D67FFFEDA60000, 37, 0xFFF7EEE000000000, 43,
                                 6364136223846793005)
    mask = (1 << 64) - 1
    mt = [seed & mask]
    for i in range(1, n):
        mt.append((f * (mt[-1] ^ (mt[-1] >> (w - 2))) + i) & mask)
    idx = n
    lower, upper = (1 << r) - 1, mask ^ ((1 << r) - 1)
    while True:
        if idx >= n:
            for i in range(n):
                y = (mt[i] & upper) | (mt[(i + 1) % n] & lower)
                mt[i] = mt[(i + m) % n] ^ (y >> 1) ^ (a if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= (y >> u) & d
        y ^= (y << s) & b
        y ^= (y << t) & c
        y ^= y >> l
        yield y & mask


def _normals(seed, count):
    """std::normal_distribution<double>(0, 1) over std::mt19937_64 as the GNU
    C++ library draws it: polar method on 2 generate_canonical() - 1 pairs,
    returning y * mult and keeping x * mult for the next call."""
    g = _mt19937_64(seed)
    out, saved = [], None
    while len(out) < count:
        if saved is not None:
            out.append(saved)
            saved = None
            continue
        while True:
            x = 2.0 * (float(next(g)) / 18446744073709551616.0) - 1.0
            y = 2.0 * (float(next(g)) / 18446744073709551616.0) - 1.0
            r2 = x * x + y * y
            if r2 <= 1.0 and r2 != 0.0:
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        saved = x * mult
        out.append(y * mult)
    return np.array(out)


def test_mt19937_64_restatement_is_the_standard_engine():
    # [rand.predef]: the 10000th output of a default-constructed mt19937_64
    g = _mt19937_64(5489)
    for _ in range(9999):
        next(g)
    assert next(g) == 9981545732273789042


@pytest.mark.parametrize("seed", [1, 2, 17])
def test_perturb_matches_the_standard_library_draws(seed):
    ref = np.sin(0.37 * np.arange(343) + 0.5) * 0.8
    out = H.perturb(ref, 0.05, seed)
    d = _normals(seed * 7919 + 13, ref.size)
    scale = 0.05 * math.sqrt(float(np.sum(ref * ref))) / math.sqrt(float(np.sum(d * d)))
    np.testing.assert_allclose(out, ref + scale * d, rtol=0, atol=1e-15)
    assert abs(np.linalg.norm(out - ref) - 0.05 * np.linalg.norm(ref)) < 1e-14
    assert not np.array_equal(out, H.perturb(ref, 0.05, seed + 1))
    assert np.array_equal(out, H.perturb(ref, 0.05, seed))


def test_perturb_rejects_zero_and_bad_rho():
    with pytest.raises(H.NumaError):
        H.perturb(np.zeros(8), 0.1, 1)
    with pytest.raises(H.NumaError):
        H.perturb(np.ones(8), 0.0, 1)
