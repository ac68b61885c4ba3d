"""CPU oracle of the RustSpice DC/TRAN hot path (test infrastructure only)."""
