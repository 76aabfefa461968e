"""MI355X-native MESSI likelihood hot path (see the package modules)."""
