"""Host-side guards that need no GPU: graph capture refuses a scattering model
that synchronises with the host, and the learned stand-in refuses in-place
evaluation."""
import types

import pytest


def test_graph_capture_refuses_host_synchronising_model():
    from frfs_amd.surrogate import DGFSVHSGLLFNOScatteringModel
    from frfs_amd.system import StdTVDRK2DGFSStepper
    sm = object.__new__(DGFSVHSGLLFNOScatteringModel)
    system = types.SimpleNamespace(sm=sm, backend=types.SimpleNamespace(torch=None))
    stp = StdTVDRK2DGFSStepper(system, graph=True)
    with pytest.raises(ValueError, match='cannot be captured'):
        stp.step(0.0, 0.01)


def test_spectral_models_are_capture_safe():
    from frfs_amd.scattering import (DGFSBiVSSGLLScatteringModel, DGFSVHSGLLNodalScatteringModel,
                                     DGFSVHSGLLScatteringModel)
    for cls in (DGFSVHSGLLScatteringModel, DGFSVHSGLLNodalScatteringModel, DGFSBiVSSGLLScatteringModel):
        assert cls.capture_safe


def test_stand_in_refuses_in_place():
    from frfs_amd.surrogate import DGFSVHSGLLFNOScatteringModel
    sm = object.__new__(DGFSVHSGLLFNOScatteringModel)
    arr = types.SimpleNamespace(_as_parameter_=12345)
    with pytest.raises(ValueError, match='alias'):
        sm.fs_batch(arr, arr)
