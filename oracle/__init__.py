"""TEST INFRASTRUCTURE ONLY: CPU oracle for the sat_amd parity tests."""
