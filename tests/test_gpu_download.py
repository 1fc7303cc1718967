"""Device-to-host copies land whole and in place.

Pageable destinations of 4 MiB and more go through the context's pinned
staging in 16-MiB chunks, each copied out by four host threads while the
next one transfers; pinned destinations (hipHostMalloc) and smaller copies
take one DMA.  Every size here is checked byte for byte, including the bytes
just outside the destination, at offsets that leave it unaligned.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import dpa_amd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def roundtrip(ctx, out, seed):
    src = np.random.default_rng(seed).integers(0, 256, out.nbytes, dtype=np.uint8)
    d = ctx.to_device(src)
    try:
        ctx.to_host(d, out)
    finally:
        ctx.free(d)
    return src


@pytest.mark.parametrize("nbytes,shift", [(1000, 0), ((4 << 20) - 1, 5), (4 << 20, 0),
                                          ((16 << 20) + 64, 0), ((37 << 20) + 13, 3),
                                          (48 << 20, 1)])
def test_copy_to_host_pageable(ctx, nbytes, shift):
    say(f"pageable download of {nbytes} bytes at offset {shift}")
    buf = np.full(nbytes + shift + 64, 0xAB, np.uint8)
    out = buf[shift:shift + nbytes]
    src = roundtrip(ctx, out, nbytes)
    assert np.array_equal(out, src)
    assert (buf[:shift] == 0xAB).all() and (buf[shift + nbytes:] == 0xAB).all()


def test_copy_to_host_pinned(ctx):
    hip = C.CDLL("libamdhip64.so")
    nbytes = (20 << 20) + 7
    p = C.c_void_p()
    assert hip.hipHostMalloc(C.byref(p), C.c_size_t(nbytes + 64), C.c_uint(0)) == 0
    try:
        say(f"pinned download of {nbytes} bytes")
        buf = np.ctypeslib.as_array((C.c_uint8 * (nbytes + 64)).from_address(p.value))
        buf[:] = 0xAB
        out = buf[:nbytes]
        src = roundtrip(ctx, out, 7)
        assert np.array_equal(out, src)
        assert (buf[nbytes:] == 0xAB).all()
    finally:
        hip.hipHostFree(p)


def test_ftle_lattice_into_reused_volume(ctx):
    """dpa's FTLE stage on a 64 x 128 x 128 field (4 MiB volume, the staged
    path) writes the same volume into a caller's array, twice in a row, as
    into a fresh one."""
    from dpa_amd import workloads as W
    shape = (64, 128, 128)
    h = np.float32(2 * np.pi / 127)
    field, sp = W.abc_field(shape, spacing=(h, h, h))
    fld = ctx.field(field, (0, 0, 0), sp)
    try:
        say("ftle_lattice 64x128x128 into a fresh and a reused array")
        ref, steps = ctx.ftle_lattice(fld, (1, 1, 1), 20, np.float32(0.02))
        vol = np.full(shape, np.nan, np.float32)
        for _ in range(2):
            got, s2 = ctx.ftle_lattice(fld, (1, 1, 1), 20, np.float32(0.02), out=vol)
            assert got is vol and s2 == steps
            assert np.array_equal(vol.view(np.uint32), ref.view(np.uint32))
        with pytest.raises(ValueError):
            ctx.ftle_lattice(fld, (1, 1, 1), 20, np.float32(0.02),
                             out=np.empty((64, 128, 127), np.float32))
    finally:
        fld.release()
