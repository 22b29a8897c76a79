"""MI355X engine for the binned BLML fit (drop-in for blml_binned.BLMLBinned's fit path)."""
from .binned import BLMLBinned, FitBatch, fit_many
from .empirical import BLMLEmpirical, empirical_many
from .modes import make_modes, parse_modes
from .quick import BLMLQuick
from .select import fc_sweep, select_fc
from .streaming import BLMLBinnedStreaming

__all__ = ['BLMLBinned', 'BLMLEmpirical', 'BLMLBinnedStreaming', 'BLMLQuick', 'FitBatch', 'fc_sweep', 'fit_many', 'empirical_many', 'make_modes', 'parse_modes', 'select_fc']
