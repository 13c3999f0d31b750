"""torchcor.core.visualize: the VTK frame writers (VTK3D, VTK3DSurface) of
torchcor/core/visualize.py, written without pyvista."""
from torchcor_amd.core.vtk import VTK3D, VTK3DSurface  # noqa: F401
