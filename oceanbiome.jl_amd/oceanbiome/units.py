"""Oceananigans.Units: time in seconds as Float64."""
second = seconds = 1.0
minute = minutes = 60 * second
hour = hours = 60 * minute
day = days = 24 * hour
year = years = 365 * day
