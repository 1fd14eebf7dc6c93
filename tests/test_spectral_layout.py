"""The algebra the frequency-domain shape of the grouped product relies on
(hpsim_amd/csrc/hps_field_fft.hip), checked on the CPU against the direct
Toeplitz sum of the direct shapes:

* the offset table S[rp][rs][comp][o] (o = zp - zs + nz - 1) laid out by
  offset modulo L = 2 nz gives a circular convolution without wrap-around
  for zp = 0 .. nz;
* two real rows transformed as one complex row are separated by Hermitian
  symmetry (E_r and E_z of a ring pair; rings rs and rs + 1 of the charges);
* the inverse of two Hermitian half spectra as one complex transform of the
  conjugated sum, conjugated back.

The transforms here are numpy's; the kernels' radix-2 loop structure is
checked against the DFT in test_radix2_loop_matches_dft below."""
import numpy as np


def _direct(S, q, nr, nz):
    W = np.zeros((nr + 1, 2, nz + 1))
    for rp in range(nr + 1):
        for c in range(2):
            for zp in range(nz + 1):
                W[rp, c, zp] = sum(S[rp, rs, c, zp - zs + nz - 1] * q[rs, zs]
                                   for rs in range(nr) for zs in range(nz))
    return W


def _split(X, L, f):
    u, v = X[f], X[(L - f) % L]
    a = complex(0.5 * (u.real + v.real), 0.5 * (u.imag - v.imag))
    b = complex(0.5 * (u.imag + v.imag), -0.5 * (u.real - v.real))
    return a, b


def _spectral(S, q, nr, nz):
    L, F, R = 2 * nz, nz + 1, 2 * (nr + 1)
    shat = np.zeros((F, R, nr), complex)
    for rp in range(nr + 1):
        for rs in range(nr):
            x = np.array([complex(S[rp, rs, 0, o], S[rp, rs, 1, o])
                          for o in [k + nz - 1 if k <= nz else k - nz - 1 for k in range(L)]])
            X = np.fft.fft(x)
            for f in range(F):
                shat[f, 2 * rp, rs], shat[f, 2 * rp + 1, rs] = _split(X, L, f)
    qhat = np.zeros((F, nr), complex)
    for rs in range(0, nr, 2):
        x = np.zeros(L, complex)
        x[:nz] = q[rs] + 1j * q[rs + 1]
        X = np.fft.fft(x)
        for f in range(F):
            qhat[f, rs], qhat[f, rs + 1] = _split(X, L, f)
    what = np.einsum("frk,fk->fr", shat, qhat)
    out = np.zeros((nr + 1, 2, nz + 1))
    for rp in range(nr + 1):
        x = np.zeros(L, complex)
        for k in range(L):
            fk = k if k <= nz else L - k
            a, b = what[fk, 2 * rp], what[fk, 2 * rp + 1]
            if k > nz:
                a, b = a.conjugate(), b.conjugate()
            x[k] = complex(a.real - b.imag, -(a.imag + b.real))
        X = np.fft.fft(x)
        out[rp, 0] = X[:nz + 1].real / L
        out[rp, 1] = -X[:nz + 1].imag / L
    return out


def test_spectral_product_equals_direct_toeplitz_sum():
    rng = np.random.default_rng(3)
    for nr, nz in ((2, 4), (4, 8), (6, 16)):
        S = rng.standard_normal((nr + 1, nr, 2, 2 * nz))
        q = rng.standard_normal((nr, nz))
        d = _direct(S, q, nr, nz)
        s = _spectral(S, q, nr, nz)
        assert np.abs(s - d).max() < 1e-12 * max(1.0, np.abs(d).max()), (nr, nz)


def test_a_wrong_offset_layout_is_caught():
    """The layout matters: offsets mapped without the modulo-L wrap give a
    different (wrong) product, so the check above can fail."""
    rng = np.random.default_rng(4)
    nr, nz = 4, 8
    S = rng.standard_normal((nr + 1, nr, 2, 2 * nz))
    q = rng.standard_normal((nr, nz))
    d = _direct(S, q, nr, nz)
    Sbad = np.roll(S, 1, axis=3)
    assert np.abs(_spectral(Sbad, q, nr, nz) - d).max() > 1e-3


def _radix2(x, L):
    """The loop structure of fft_rows (bit reversal, then radix-2 stages
    with twiddles exp(-2 pi i k / L) at stride L / (2 half))."""
    lg = L.bit_length() - 1
    x = x.astype(complex).copy()
    for k in range(L):
        j = int(format(k, f"0{lg}b")[::-1], 2)
        if j > k:
            x[k], x[j] = x[j], x[k]
    tw = np.exp(-2j * np.pi * np.arange(L // 2) / L)
    half = 1
    while half < L:
        step = L // (2 * half)
        for j in range(L // 2):
            g, k = divmod(j, half)
            i0 = g * 2 * half + k
            i1 = i0 + half
            b = x[i1] * tw[k * step]
            a = x[i0]
            x[i0], x[i1] = a + b, a - b
        half <<= 1
    return x


def test_radix2_loop_matches_dft():
    rng = np.random.default_rng(5)
    for L in (8, 16, 64, 256):
        x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
        assert np.abs(_radix2(x, L) - np.fft.fft(x)).max() < 1e-11
