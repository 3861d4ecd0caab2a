"""TEST INFRASTRUCTURE ONLY: CPU oracle for the MHD hot path (see oracle.py, oracle.cc)."""
