"""Adapters from other codes' input formats (reference `adapters/__init__.py`)."""
