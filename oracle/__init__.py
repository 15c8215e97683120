"""CPU oracle of the segmented sort (test infrastructure; never imported by the product)."""
