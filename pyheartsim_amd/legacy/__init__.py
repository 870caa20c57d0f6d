"""Device versions of the reference's legacy/ analysis helpers."""
