"""MI355X-native linear-solve and stamping core for RustSpice's Newton loop."""
