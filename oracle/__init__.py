"""CPU oracle of the H.264 analysis stage: test infrastructure only (the checker, never the product)."""
