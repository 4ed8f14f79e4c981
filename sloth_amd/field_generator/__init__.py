"""Random density fields (mirrors the reference's field_generator package)."""
