"""CPU oracle of the threeD triangle count -- test infrastructure only (the
checker and the CPU baseline); never imported by the product package."""
