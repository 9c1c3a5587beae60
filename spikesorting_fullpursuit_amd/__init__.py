"""MI355X-native full binary pursuit (drop-in for the reference's
spikesorting_fullpursuit.overlap.binary_pursuit_parallel.binary_pursuit)."""
__version__ = "0.1.0"
