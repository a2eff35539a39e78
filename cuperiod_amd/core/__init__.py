"""Data model of the BLS engine: light curves, grids, settings, results, errors."""
