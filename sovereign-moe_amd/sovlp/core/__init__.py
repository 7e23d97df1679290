"""Boundary types and the device backend."""
