"""Drop-in name: ``import xgrid`` resolves to the MI355X backend package."""
from xgrid_amd import *  # noqa: F401,F403
from xgrid_amd import __all__  # noqa: F401
