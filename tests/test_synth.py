"""The seeded synthetic volume source (the workloads' stand-in for a brick
store) keeps its bytes: whole volumes and on-demand bricks of every level,
u8 and u16, hashed against a digest recorded from the scalar implementation
of the voxel loop.  The library runs whichever of its AVX-512 / AVX2 /
x86-64 builds the host CPU supports, so this also pins the vector builds to
the scalar one."""
import hashlib

import numpy as np
import pytest

SCALAR_DIGEST = "cbfd355f2f10f68d1fabeaa4dc407b8793b83376c4470fe4fc5ce481c674304f"


def test_synthetic_bytes_match_scalar_digest(gcv):
    G = gcv
    h = hashlib.sha256()
    for nb, bytes_, dims in ((1, 1, (256, 256, 256)), (2, 2, (300, 200, 100))):
        p = G.synth_params(7 + nb, 12, sigma=(8.0, 40.0), amp=(60.0, 255.0) if bytes_ == 1 else (15000.0, 65535.0),
                           noise=8 if bytes_ == 1 else 2048, threads=1)
        h.update(G.synth_volume(p, dims, bytes_).tobytes())
        vc = G.volume_cfg(dims, 32, 1, bytes_)
        src = G.Source.synthetic(vc, p)
        for lv in range(vc.nb_levels):
            st, out = src.get(vc, lv, [0, 1, 0] if lv == 0 else [0, 0, 0])
            h.update(out.tobytes())
    assert h.hexdigest() == SCALAR_DIGEST


@pytest.mark.parametrize("threads", [1, 4])
def test_synthetic_volume_independent_of_threads(gcv, threads):
    G = gcv
    p1 = G.synth_params(11, 7, sigma=(4.0, 12.0), amp=(60.0, 255.0), noise=8, threads=1)
    pn = G.synth_params(11, 7, sigma=(4.0, 12.0), amp=(60.0, 255.0), noise=8, threads=threads)
    a = G.synth_volume(p1, (70, 33, 19), 1)
    b = G.synth_volume(pn, (70, 33, 19), 1)
    assert np.array_equal(a, b)
