"""Host file formats: the C++ ConfigPack pair and the four-column function table."""
