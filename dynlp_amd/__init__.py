"""MI355X-native hot path of dynamic-graph harmonic label propagation (DynLP+)."""
__version__ = "0.1.0"
