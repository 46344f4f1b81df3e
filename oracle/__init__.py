"""CPU oracle of the NonhydrostaticModel time step (test infrastructure only)."""
