"""Pins of the oracle's RNG and camera (include/cuda/utils/lcg_rng.cuh, camera_raygen)."""
import numpy as np

from conftest import camera_rays

M32 = 0xFFFFFFFF


def murmur3_x86_32(orc, key: bytes, seed: int) -> int:
    """MurmurHash3_x86_32 composed from the oracle's block mix and finaliser (tail-free keys)."""
    assert len(key) % 4 == 0
    h = seed
    for i in range(0, len(key), 4):
        h = orc.murmur_mix(h, int.from_bytes(key[i:i + 4], "little"))
    return orc.murmur_fmix(h ^ len(key))


def test_murmur3_published_vectors(orc):
    # Published MurmurHash3_x86_32 verification values (SMHasher test vectors).
    assert murmur3_x86_32(orc, b"", 0) == 0
    assert murmur3_x86_32(orc, b"", 1) == 0x514E28B7
    assert murmur3_x86_32(orc, b"", 0xFFFFFFFF) == 0x81F16F39
    assert murmur3_x86_32(orc, b"\x00\x00\x00\x00", 0) == 0x2362F9DE
    assert murmur3_x86_32(orc, b"\xff\xff\xff\xff", 0) == 0x76293B50
    assert murmur3_x86_32(orc, b"\x21\x43\x65\x87", 0) == 0xF55B516B
    assert murmur3_x86_32(orc, b"\x21\x43\x65\x87", 0x5082EDEE) == 0x2362F9DE
    assert murmur3_x86_32(orc, b"aaaa", 0x9747B28C) == 0x5A97808A


def test_lcg_seed_composition(orc):
    # get_lcg_rng: mix(mix(0, linear), 512*frame + index), then fmix (lcg_rng.cuh:65-72).
    for idx, frame, lin in [(0, 0, 0), (3, 7, 12345), (511, 1, 0xFFFFFFFF)]:
        expect = orc.murmur_fmix(orc.murmur_mix(orc.murmur_mix(0, lin), (512 * frame + idx) & M32))
        assert orc.lcg_seed(idx, frame, lin) == expect


def test_lcg_closed_form(orc):
    # state_{k+1} = 1664525 * state_k + 1013904223 mod 2^32; float = ldexp((float)state, -32).
    s = 0x12345678
    got = orc.lcg_floats(s, 6)
    expect = []
    for _ in range(6):
        s = (1664525 * s + 1013904223) & M32
        expect.append(np.float32(np.float32(s) * np.float32(2.0 ** -32)))
    assert np.array_equal(got, np.array(expect, np.float32))


def test_raygen_matches_numpy_restatement(nb, orc):
    w, h = 17, 9
    cam, rays = camera_rays(nb, orc, w, h, sample=2, frame=5)
    c = nb.camera_array(cam)
    for i in [0, 5, 77, w * h - 1]:
        s = orc.lcg_seed(2, 5, i)
        u = orc.lcg_floats(s, 2)
        xf = np.float32(np.float32(i % w) + u[0]) / np.float32(w)
        yf = np.float32(np.float32(i // w) + u[1]) / np.float32(h)
        d = (xf * c[3:6] + yf * c[6:9]) + c[9:12]
        d = d.astype(np.float32)
        n = np.float32(1.0) / np.sqrt(np.float32((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
        assert np.array_equal(rays["d"][i], (d * n).astype(np.float32))
        assert np.array_equal(rays["o"][i], c[0:3])
    # Centre pixel direction points at the target.
    assert np.allclose(rays["d"][(h // 2) * w + w // 2], [0, 0, -1], atol=0.08)


def test_expf_against_numpy(orc):
    x = np.concatenate([np.linspace(-87, 88, 20001, dtype=np.float32),
                        np.float32([0.0, -0.0, 1e-8, -1e-8, 0.5, -0.34657359, 0.34657359])])
    got = orc.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(np.float32(ref.astype(np.float32))).astype(np.float64)
    err = np.abs(got - ref) / ulp
    assert err.max() <= 2.0, err.max()
    assert orc.expf(np.float32([200.0]))[0] == np.inf
    assert orc.expf(np.float32([-200.0]))[0] == 0.0
