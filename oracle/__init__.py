"""CPU oracle of the volumetric merging estimators (test infrastructure only)."""
