"""The oracle: test infrastructure only (see orc.h). Not imported by espice_amd."""
