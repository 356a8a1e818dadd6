"""Host parts of the resident sampler (CPU only): the Philox4x32-10 generator
against its published known-answer vectors (Random123 kat_vectors), and the
reduced A-matrix of states with closed forms."""
import ctypes

import numpy as np


def philox(ctr, key):
    from gbskernels_amd import _lib
    c = (ctypes.c_uint * 4)(*ctr)
    k = (ctypes.c_uint * 2)(*key)
    o = (ctypes.c_uint * 4)()
    _lib.lib().gbs_philox4x32_10(ctypes.addressof(c), ctypes.addressof(k), ctypes.addressof(o))
    return list(o)


def test_philox_known_answers():
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_reduced_A_closed_forms():
    from gbskernels_amd.sampling import reduced_A
    # vacuum: A = 0
    assert np.allclose(reduced_A(np.eye(6), 2), 0)
    # single-mode squeezed vacuum: A = B (+) B*, |B| = tanh r
    r = 0.7
    cov = np.diag([np.exp(-2 * r), np.exp(2 * r)])
    A = reduced_A(cov, 1)
    assert abs(abs(A[0, 0]) - np.tanh(r)) < 1e-14 and abs(A[0, 1]) < 1e-14
    assert abs(A[1, 1] - np.conj(A[0, 0])) < 1e-14
    # tracing out a vacuum mode leaves the squeezed mode's A unchanged
    cov2 = np.diag([np.exp(-2 * r), 1.0, np.exp(2 * r), 1.0])
    assert np.allclose(reduced_A(cov2, 1), A)
