"""Drop-in replacements for the reference driver's hot-path modules.

Put this directory on sys.path ahead of the reference's LIBRARIES/ (or copy
the three modules there) and MWDiem.py runs unchanged:
  gpu_fine_contact_detection.Contact_detection_GPU
  gpu_Verlet.Verlet_GPU
  LIBRARY_Coarse_contact_detection.Linked_cell_hash
"""
