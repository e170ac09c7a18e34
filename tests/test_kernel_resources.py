"""The compiled kernels' resources, from hipcc's kernel-resource-usage report
that the build writes beside each object: no kernel uses scratch memory, and
the path tracer's shading kernels keep four waves per SIMD.

Scratch in a shading kernel costs every scene, not only the scenes that reach
the code that needs it: when a helper that takes the kernel's argument block
by reference is compiled as a call, the kernel copies the whole block to
private memory at entry (the material node-graph interpreter once made the
inliner give up on such helpers, and the fused bounce ran at a third of its
speed)."""
import pytest

from pine_amd import build as B


def report(src):
    r = B.kernel_resources(src)
    if r is None:
        pytest.skip(f"no resource report for {src}: the library was not built here")
    assert r, f"empty resource report for {src}"
    return r


@pytest.mark.parametrize("src", ["accel.hip", "render.hip"])
def test_no_traversal_or_rendering_kernel_uses_scratch(src):
    r = report(src)
    scratch = {k: v.get("ScratchSize") for k, v in r.items() if v.get("ScratchSize", 0) != 0}
    assert not scratch, scratch


def test_guide_scratch_is_only_the_training_kernels_spills():
    """The gradient kernels hold a layer's activations and their gradients at
    once (up to 256 VGPRs) and spill a little; inference and Adam use none."""
    r = report("guide.hip")
    scratch = {k: v.get("ScratchSize") for k, v in r.items() if v.get("ScratchSize", 0) != 0}
    assert all("k_guide_grad" in k or "k_guide_apply" in k for k in scratch), scratch
    assert all(v <= 256 for v in scratch.values()), scratch


def test_shading_kernels_keep_four_waves_per_simd():
    r = report("render.hip")
    shade = {k: v for k, v in r.items() if "k_shadeILb" in k or "k_shade_spec" in k}
    assert len(shade) == 3, sorted(shade)
    low = {k: v.get("Occupancy") for k, v in shade.items() if v.get("Occupancy", 0) < 4}
    assert not low, low
