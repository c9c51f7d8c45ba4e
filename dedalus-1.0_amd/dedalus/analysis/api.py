"""Public analysis API (dedalus/analysis/api.py)."""

from .analysis_set import AnalysisSet, Snapshot, TrackMode, VolumeAverage, PowerSpectrum
from .volume_average import VolumeAverageSet
