"""``bmmpy.candidate``: the search result record (bmmpy/candidate.py)."""
from bmmpy_amd._bmmpy import Candidate

__all__ = ["Candidate"]
