"""CPU oracle of RecView's FBP path -- test infrastructure only (see oracle.h)."""
