"""Ragged (mixed-size) batches through the persistent work-queue kernel."""
from __future__ import annotations


class RaggedBatch:  # filled in with the ragged kernel
    def __init__(self, *a, **k):
        raise NotImplementedError("ragged path not built yet")
