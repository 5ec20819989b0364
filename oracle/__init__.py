"""CPU oracle of DIABLO's channel path: test infrastructure, not product."""
