"""pyqed.HEOM (pyqed/HEOM/heom.py): the second copy of the legacy single-exponential HEOM solver."""
