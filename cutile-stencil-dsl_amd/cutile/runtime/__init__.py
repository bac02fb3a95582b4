"""package"""
