"""MI355X (gfx950) HIP platform for HiFlow3's sparse linear-algebra hot path.

The directory name is not a Python identifier, so callers load the package
by path and register it as `hiflow3_amd`:

    spec = importlib.util.spec_from_file_location(
        "hiflow3_amd", "<repo>/hiflow3_2.0_amd/__init__.py",
        submodule_search_locations=["<repo>/hiflow3_2.0_amd"])
"""
from . import hfh  # noqa: F401
from .hfh import *  # noqa: F401,F403
