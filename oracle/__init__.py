"""CPU oracle of the MD step -- test infrastructure, never the product path."""
