"""CPU oracle of the MHIT36 timestep (test infrastructure only)."""
