"""Ensemble analysis: isotropic S(k) and pair statistics (GPU moments, host binning)."""
