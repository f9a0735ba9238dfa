"""The numpy restatement of smooth2's sub-point count for several droplets
(oracle_ffi.sampled_cc), which checks the GPU's multi-droplet initial
condition, reproduces the CPU restatement's smooth2 (main.f:5678-5752,
itself bit-exact against the flang-built reference) for one droplet."""
import numpy as np

from oracle_ffi import Oracle, inputdata_text, sampled_cc


def test_sampler_reproduces_smooth2():
    for kw in (dict(nx=18, ny=18, nz=18), dict(nx=34, ny=18, nz=18, y0=0.002, radius=0.0005)):
        o = Oracle(inputdata_text(**kw))
        cc = o.field("cc").copy()
        h = o.st.h
        drops = [((kw["nx"] - 2) * h / 4.0, o.st.y0, o.st.z0, o.st.radius, 0.0, 0.0, 0.0)]
        want = sampled_cc(cc.shape, h, drops)
        o.close()
        assert np.array_equal(cc[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1])
        assert 0.0 < want.sum()
