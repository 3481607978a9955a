"""MI355X-native phonon Monte Carlo trajectory engine (montecarlocpp path)."""
