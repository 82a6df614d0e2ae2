"""The reference's module name: ``import snapy`` gives this build's
MeshBlockOptions, MeshBlock, StageExecutor and the rest of the surface of
``snapy_amd`` (python/snapy/__init__.py of the reference exports the same
names from its pybind module)."""

from snapy_amd import *  # noqa: F401,F403
from snapy_amd import __all__  # noqa: F401
