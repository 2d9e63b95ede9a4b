"""Dataset readers (slamtb/data/__init__.py): the FTB frame format and TUM
trajectory files."""
from .ftb import FTBDataset, load_ftb, load_trajectory, write_ftb
