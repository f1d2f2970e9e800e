"""MI355X-native k-NN search path: exact FP32 scan, BF16/FP8/SQ8/SQ4 candidate
scans with FP32 re-score, top-k selection, behind the competition plugin
surface (``Algorithm``) and a plain C ABI (``include/annmi.h``)."""
from .algorithm import Algorithm  # noqa: F401
from .index import Index  # noqa: F401

__all__ = ["Algorithm", "Index"]
