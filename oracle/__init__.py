"""The CPU checkers of the relaxation (test infrastructure, not the product)."""
