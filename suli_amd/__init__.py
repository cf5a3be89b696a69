"""MI355X-native pressure-projection solver for SuLi (graddin + posdin)."""
