"""pyqed.deom: the reference's alias package for the HEOM solver (pyqed/deom/__init__.py re-exports its
copy of heom/deom.py, which scripts such as heom/propagator.py:7-10 import from).  Here it re-exports
pyqed_amd.heom.deom; pyqed_amd.deom.heom is the legacy solver module of pyqed.HEOM.heom, which the
reference's pyqed/deom/heom.py duplicates byte for byte."""
from ..heom.deom import *  # noqa: F401,F403
from ..heom.deom import __all__  # noqa: F401
