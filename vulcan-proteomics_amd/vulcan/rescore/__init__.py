"""Rescoring-side modules: the learned candidate prefilter."""
