"""Host-side utilities around the simulators."""
