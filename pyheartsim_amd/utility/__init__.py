"""Device versions of the reference's post-processing helpers (utility/)."""
