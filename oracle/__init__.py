"""CPU oracle of the EmptyRoom grid / Widom / pocket arithmetic (test infrastructure only)."""
